"""The synthetic workload generators (CPU): the FFT-summed Fourier modes equal
the direct cosine sums, and the masks / seeds follow the stated recipes."""
import numpy as np
import pytest

import csipm.workloads as wl


def _direct(dims, nmodes, seed, kmax=None):
    rng = np.random.default_rng(seed)
    d = len(dims)
    coords = wl._grid_coords(dims)
    if d == 1:
        hi = kmax if kmax is not None else dims[0] // 2 - 1
        freqs = rng.choice(np.arange(1, hi + 1), size=nmodes, replace=False)[:, None]
    else:
        k = kmax if kmax is not None else 127
        freqs = rng.integers(-k, k + 1, size=(nmodes, d))
    amps = rng.standard_normal(nmodes)
    phases = rng.uniform(0.0, 2 * np.pi, nmodes)
    x = np.zeros(int(np.prod(dims)))
    for j in range(nmodes):
        arg = sum(freqs[j, ax] * coords[ax] / dims[ax] for ax in range(d))
        x += amps[j] * np.cos(2 * np.pi * arg + phases[j])
    return x


@pytest.mark.parametrize("dims,nmodes,kmax", [((1024,), 12, 511), ((63, 45), 40, 20),
                                              ((15, 12, 9), 30, 4)])
def test_signal_fft_equals_cosine_sum(dims, nmodes, kmax):
    got = wl._signal(np.random.default_rng(7), dims, nmodes, kmax=kmax)
    want = _direct(dims, nmodes, 7, kmax=kmax)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_config_shapes_and_masks():
    p = wl.punched_2d_255()
    assert p.dims == (255, 255) and p.lam == 0.05
    assert 0.08 < len(p.missing) / p.N < 0.14        # r=3 discs on a 16-lattice
    assert np.all(p.z0[p.missing] == 0.0)
    shapes = wl.uneven_shapes()
    assert len(shapes) == 384
    n1 = [s[0] for s in shapes if len(s) == 1]
    assert min(n1) >= 512 and max(n1) <= 65536
    assert all(s[0] % 2 == 1 and 63 <= s[0] <= 511 for s in shapes if len(s) == 2)
    arr, sh, lams = wl.stream_requests(nreq=2000)
    assert np.all(np.diff(arr) >= 0)
    frac = np.array([len(s) for s in sh])
    assert abs((frac == 1).mean() - 0.70) < 0.04 and abs((frac == 3).mean() - 0.05) < 0.02
    q = wl.stream_problem((127, 127), seed=3)
    assert q.lam == 0.05 and np.all(q.z0[q.missing] == 0.0)
