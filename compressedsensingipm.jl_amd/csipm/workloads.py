"""Synthetic signals, punch masks and problem instances.

Two families:

* the reference's own test generators, restated:
  - `example_signal` : test/ipm_example_1D.jl:6-9, ipm_example_2D.jl:7, ipm_example_3D.jl:8
  - `centering` / `punching_optimized` : test/punching_centering.jl:74-114, 131-201
* the benchmark workloads (seeded numpy PCG64 streams, seed = 1_000_003*family
  + instance, families 0..4 in the order below; draw order modes ->
  amplitudes -> phases -> noise -> structured mask -> random dropout):
  `signal_1d_1024`, `punched_2d_255`, `punched_3d`, `uneven_batch`,
  `stream_problem` / `stream_requests`.

All arrays are flat vectors in Julia column-major order (dimension 1 fastest);
`missing` holds 0-based linear indices and z0 is zero there, as
punching_centering.jl:8 zeroes data[index_missing].
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Problem:
    dims: tuple
    z0: np.ndarray          # observed samples (zero at missing), length N
    lam: float
    missing: np.ndarray     # 0-based linear indices, sorted, unique
    name: str = ""
    x_true: np.ndarray | None = None

    @property
    def N(self):
        return int(np.prod(self.dims))


def _grid_coords(dims):
    """Coordinates (n1, n2, n3) of every sample in Julia linear order."""
    axes = [np.arange(n) for n in dims]
    mesh = np.meshgrid(*axes, indexing="ij")
    return [m.reshape(-1, order="F") for m in mesh]


# ---------------------------------------------------------------------------
# reference test generators
# ---------------------------------------------------------------------------

def example_signal(dims):
    """Noise-free signals of the reference's IPM examples."""
    dims = tuple(dims)
    if len(dims) == 1:
        n = dims[0]
        t = np.arange(n)
        x1 = 2 * np.cos(2 * np.pi * t * 6 / n) + 3 * np.sin(2 * np.pi * t * 6 / n)
        x2 = 4 * np.cos(2 * np.pi * t * 10 / n) + 5 * np.sin(2 * np.pi * t * 10 / n)
        x3 = 6 * np.cos(2 * np.pi * t * 40 / n) + 7 * np.sin(2 * np.pi * t * 40 / n)
        return x1 + x2 + x3
    if len(dims) == 2:
        nt, ns = dims
        t = np.arange(nt)
        s = np.arange(ns)
        a = np.cos(2 * np.pi * 2 / nt * t) + 2 * np.sin(2 * np.pi * 2 / nt * t)
        b = np.cos(2 * np.pi * 3 / ns * s) + 2 * np.sin(2 * np.pi * 3 / ns * s)
        return np.outer(a, b).reshape(-1, order="F")
    n1, n2, n3 = dims
    i, j, k = np.arange(n1), np.arange(n2), np.arange(n3)
    a = np.cos(2 * np.pi * 1 / n1 * i) + 2 * np.sin(2 * np.pi * 1 / n1 * i)
    b = np.cos(2 * np.pi * 2 / n2 * j) + 2 * np.sin(2 * np.pi * 2 / n2 * j)
    c = np.cos(2 * np.pi * 3 / n3 * k) + 2 * np.sin(2 * np.pi * 3 / n3 * k)
    return np.einsum("i,j,k->ijk", a, b, c).reshape(-1, order="F")


def centering(dims, missing_prob):
    """Lattice centres (1-based), punching_centering.jl:84-114."""
    dims = tuple(dims)
    n = int(np.prod(dims))
    d = len(dims)
    denom = {1: 3, 2: 5, 3: 7}[d]
    per = (n * missing_prob / denom) ** (1.0 / d)
    axes = [range(1, nd + 1, int(math.ceil(nd / per))) for nd in dims]
    mesh = np.meshgrid(*[np.asarray(list(a)) for a in axes], indexing="ij")
    return np.stack([m.reshape(-1, order="F") for m in mesh], axis=1)


def punching_optimized(dims, centers):
    """Radius-1 punch (the centre and its 2d face neighbours), 0-based linear
    indices in the order punching_centering.jl:131-201 emits them."""
    dims = tuple(dims)
    d = len(dims)
    offs = [np.zeros(d, int)]
    for ax in range(d):
        for s in (-1, 1):
            o = np.zeros(d, int)
            o[ax] = s
            offs.append(o)
    # the reference loops i2, j2, k2 in increasing order and keeps |offset|_1 <= 1
    offs = np.asarray(sorted(offs, key=lambda o: tuple(o)), dtype=np.int64)
    strides = np.cumprod((1,) + dims[:-1]).astype(np.int64)
    c = np.asarray(centers, dtype=np.int64).reshape(-1, d)
    # every centre with every offset, centre-major (the loop order above)
    p = c[:, None, :] + offs[None, :, :]
    ok = np.all((p >= 1) & (p <= np.asarray(dims, dtype=np.int64)), axis=2)
    return ((p - 1) @ strides)[ok].astype(np.int64)


def reference_example(dims, check=True, seed=1):
    """Problem of test/ipm_example_{1,2,3}D.jl.  check=True: no noise, no
    missing samples, lambda = 0 (the known-answer configuration)."""
    dims = tuple(dims)
    x = example_signal(dims)
    if check:
        return Problem(dims, x.copy(), 0.0, np.zeros(0, np.int64), "example-check", x)
    rng = np.random.default_rng(seed)
    y = x + rng.standard_normal(x.size)
    missing = np.unique(punching_optimized(dims, centering(dims, 0.15)))
    y[missing] = 0.0
    lam = 1.0 if len(dims) == 1 else 5.0
    return Problem(dims, y, lam, missing, "example", x)


# ---------------------------------------------------------------------------
# benchmark workloads
# ---------------------------------------------------------------------------

def _seed(family, instance):
    return 1_000_003 * family + instance


def _signal(rng, dims, nmodes, kmax=None):
    """Sum of nmodes random Fourier modes: integer wavevectors, N(0,1)
    amplitudes, U[0, 2pi) phases.  The modes are summed by one inverse FFT of
    the sparse spectrum sum_j a_j exp(i phi_j) delta(k - f_j) (the same values
    as summing the cosines, to rounding, at N log N cost)."""
    dims = tuple(dims)
    d = len(dims)
    if d == 1:
        n = dims[0]
        hi = kmax if kmax is not None else n // 2 - 1
        freqs = rng.choice(np.arange(1, hi + 1), size=nmodes, replace=False)[:, None]
    else:
        lo = -(kmax if kmax is not None else 127)
        hi = kmax if kmax is not None else 127
        freqs = rng.integers(lo, hi + 1, size=(nmodes, d))
    amps = rng.standard_normal(nmodes)
    phases = rng.uniform(0.0, 2 * np.pi, nmodes)
    # C-order array of the reversed shape: the last axis is Julia's dimension 1
    spec = np.zeros(dims[::-1], dtype=np.complex128)
    for j in range(nmodes):
        idx = tuple(int(freqs[j, ax]) % dims[ax] for ax in reversed(range(d)))
        spec[idx] += amps[j] * np.exp(1j * phases[j])
    x = np.real(np.fft.ifftn(spec)) * spec.size
    return x.reshape(-1)


def _gaps_1d(rng, n, ngaps=3, width=32):
    starts = rng.integers(0, max(n - width, 0) + 1, size=ngaps)
    return np.concatenate([np.arange(s, min(s + width, n)) for s in starts])


def _lattice_balls(dims, spacing, radius):
    """Balls of `radius` around every lattice point spacing*(a, b, c), 0-based,
    clipped at the borders."""
    dims = tuple(dims)
    d = len(dims)
    r = int(radius)
    rng_off = np.arange(-r, r + 1)
    offs = np.stack(np.meshgrid(*([rng_off] * d), indexing="ij"), -1).reshape(-1, d)
    offs = offs[(offs ** 2).sum(1) <= radius * radius]
    centres = np.stack(np.meshgrid(*[np.arange(0, n, spacing) for n in dims], indexing="ij"),
                       -1).reshape(-1, d)
    pts = (centres[:, None, :] + offs[None, :, :]).reshape(-1, d)
    ok = np.all((pts >= 0) & (pts < np.asarray(dims)), axis=1)
    pts = pts[ok]
    strides = np.cumprod((1,) + dims[:-1])
    return pts @ strides


def _finish(dims, x, noise_sd, rng, structured, dropout_frac, lam, name):
    n = int(np.prod(dims))
    y = x + noise_sd * rng.standard_normal(n)
    miss = [np.asarray(structured, dtype=np.int64)]
    if dropout_frac > 0:
        k = int(round(dropout_frac * n))
        miss.append(rng.choice(n, size=k, replace=False))
    missing = np.unique(np.concatenate(miss)) if miss else np.zeros(0, np.int64)
    y[missing] = 0.0
    return Problem(tuple(dims), y, float(lam), missing.astype(np.int64), name, x)


def signal_1d_1024(instance=0):
    """1D N=1024, 12 modes, 3 gaps of 32 + 5% dropout, lambda 0.1."""
    rng = np.random.default_rng(_seed(0, instance))
    dims = (1024,)
    x = _signal(rng, dims, 12, kmax=511)
    noise = 0.01 * rng.standard_normal(dims[0])
    gaps = _gaps_1d(rng, dims[0])
    k = int(round(0.05 * dims[0]))
    drop = rng.choice(dims[0], size=k, replace=False)
    y = x + noise
    missing = np.unique(np.concatenate([gaps, drop])).astype(np.int64)
    y[missing] = 0.0
    return Problem(dims, y, 0.1, missing, "1d-1024", x)


def punched_2d_255(instance=0):
    """2D 255x255, 40 modes, discs r=3 on a spacing-16 lattice, lambda 0.05."""
    rng = np.random.default_rng(_seed(1, instance))
    dims = (255, 255)
    x = _signal(rng, dims, 40)
    return _finish(dims, x, 0.01, rng, _lattice_balls(dims, 16, 3), 0.0, 0.05,
                   "2d-255x255-punched")


def punched_3d(instance=0, side=255):
    """3D side^3, 200 modes, balls r=2 on a spacing-8 lattice + 1% dropout,
    lambda 0.05."""
    rng = np.random.default_rng(_seed(2, instance))
    dims = (side, side, side)
    x = _signal(rng, dims, 200, kmax=min(127, (side - 1) // 2))
    return _finish(dims, x, 0.01, rng, _lattice_balls(dims, 8, 2), 0.01, 0.05,
                   f"3d-{side}^3-punched")


def uneven_shapes(n1d=256, n2d=128):
    """Shapes of the batched workload: 1D N log-uniform in [512, 65536];
    2D odd sides uniform in [63, 511]."""
    rng = np.random.default_rng(_seed(3, 0))
    n1 = np.rint(np.exp(rng.uniform(np.log(512), np.log(65536), n1d))).astype(int)
    sides = rng.choice(np.arange(63, 512, 2), size=n2d)
    return [(int(n),) for n in n1] + [(int(s), int(s)) for s in sides]


def batch_instance(dims, seed, lam=0.1):
    """One instance of the batched / serving workloads: 10% dropout plus 3
    gaps of 32 (1D), discs of radius 3 (2D) or balls of radius 2 (3D)."""
    rng = np.random.default_rng(seed)
    dims = tuple(dims)
    if len(dims) == 1:
        n = dims[0]
        x = _signal(rng, dims, 12 + n // 128, kmax=(n - 1) // 2)
        noise = 0.01 * rng.standard_normal(n)
        structured = _gaps_1d(rng, n)
        drop = rng.choice(n, size=int(round(0.1 * n)), replace=False)
        y = x + noise
        missing = np.unique(np.concatenate([structured, drop])).astype(np.int64)
        y[missing] = 0.0
        return Problem(dims, y, lam, missing, f"1d-{n}", x)
    kmax = (min(dims) - 1) // 2
    nmodes = 40 if len(dims) == 2 else 200
    x = _signal(rng, dims, nmodes, kmax=kmax)
    noise = 0.01 * rng.standard_normal(x.size)
    spacing = 16 if len(dims) == 2 else 8
    radius = 3 if len(dims) == 2 else 2
    centres_struct = _lattice_balls(dims, spacing, radius)
    y = x + noise
    drop = rng.choice(x.size, size=int(round(0.1 * x.size)), replace=False)
    missing = np.unique(np.concatenate([centres_struct, drop])).astype(np.int64)
    y[missing] = 0.0
    return Problem(dims, y, lam, missing, "x".join(map(str, dims)), x)


def uneven_batch(n1d=256, n2d=128):
    shapes = uneven_shapes(n1d, n2d)
    return [batch_instance(s, _seed(3, i + 1)) for i, s in enumerate(shapes)]


def stream_problem(shape, seed):
    """One serving request: the 1D / 2D / 3D generator above at its shape
    (1D: 12 + N/128 modes, 3 gaps of 32 + 5% dropout, lambda 0.1; 2D: 40
    modes, r=3 discs on a 16-lattice, lambda 0.05; 3D: 200 modes, r=2 balls
    on an 8-lattice + 1% dropout, lambda 0.05)."""
    rng = np.random.default_rng(seed)
    dims = tuple(shape)
    if len(dims) == 1:
        n = dims[0]
        x = _signal(rng, dims, 12 + n // 128, kmax=(n - 1) // 2)
        return _finish(dims, x, 0.01, rng, _gaps_1d(rng, n), 0.05, 0.1, f"1d-{n}")
    if len(dims) == 2:
        x = _signal(rng, dims, 40, kmax=min(127, (min(dims) - 1) // 2))
        return _finish(dims, x, 0.01, rng, _lattice_balls(dims, 16, 3), 0.0, 0.05,
                       "x".join(map(str, dims)))
    x = _signal(rng, dims, 200, kmax=min(127, (min(dims) - 1) // 2))
    return _finish(dims, x, 0.01, rng, _lattice_balls(dims, 8, 2), 0.01, 0.05,
                   "x".join(map(str, dims)))


def stream_requests(nreq=10_000, rate=50.0):
    """Poisson arrivals (exponential gaps) and request shapes of the serving
    workload: 70% 1D (N log-uniform [1024, 65536]), 25% 2D (odd side in
    [127, 511]), 5% 3D (odd side in [63, 191])."""
    rng = np.random.default_rng(_seed(4, 0))
    gaps = rng.exponential(1.0 / rate, nreq)
    arrivals = np.cumsum(gaps)
    kind = rng.uniform(size=nreq)
    shapes = []
    for u in kind:
        if u < 0.70:
            shapes.append((int(np.rint(np.exp(rng.uniform(np.log(1024), np.log(65536))))),))
        elif u < 0.95:
            s = int(rng.choice(np.arange(127, 512, 2)))
            shapes.append((s, s))
        else:
            s = int(rng.choice(np.arange(63, 192, 2)))
            shapes.append((s, s, s))
    lams = [0.1 if len(s) == 1 else 0.05 for s in shapes]
    return arrivals, shapes, lams
