"""CPU FP64 oracle for the beta <-> DFT packing maps.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product path.

What it restates
----------------
The Hermitian packing of a real spectrum into a real vector beta of length
N = N1*N2*N3 used by CompressedSensingIPM.jl:

* 1D  src/mapping_cpu_kernels.jl:11-21   (DFT -> beta),  :363-373 (beta -> DFT)
* 2D  src/mapping_cpu_kernels.jl:38-87   (DFT -> beta),  :392-467 (beta -> DFT)
* 3D  src/mapping_cpu_kernels.jl:109-351 (DFT -> beta),  :494-738 (beta -> DFT)

The reference enumerates beta as a fixed sequence of blocks (vertices, edges,
faces, interior; real block then imaginary block).  Here that sequence is built
once as a table: for every beta index i it records the full-spectrum element
the entry represents, whether it is the real or the imaginary part, and whether
it is a self-conjugate vertex (no sqrt(2) factor).  Both maps are then read off
the table.

Odd sizes
---------
The reference only defines even sizes (src/fft_utils.jl:168,217).  Per
dimension the table uses

* self-conjugate set S = {0, N/2} (N even) or {0} (N odd),
* representatives  R = 1 .. P  and second half  H = M+1 .. M+P,
  with M = N // 2 and P = (N-1) // 2,
* conjugate partner of index a:  (N - a) mod N,

and keeps the reference's block order.  For all-even sizes this is exactly the
reference layout (pinned bit-exactly in tests against three independent
restatements of the reference); for odd sizes it is the documented extension,
pinned by invariants (exact Hermitian output, round trip, norm preservation).

Arithmetic (bit-exactness): DFT -> beta multiplies by the double sqrt(2)
(`sqrt(2) * real(v)`, `-sqrt(2) * imag(v)`); beta -> DFT divides each part by
the same double (`(a + im*b) / sqrt(2)` is evaluated per component in Julia).
"""
from __future__ import annotations

import numpy as np

SQRT2 = float(np.sqrt(np.float64(2.0)))  # identical to Julia's sqrt(2)


def dim_info(n: int):
    """(M, self_set, P) for one dimension of length n."""
    m = n // 2
    selfs = [0, m] if n % 2 == 0 else [0]
    p = (n - 1) // 2
    return m, selfs, p


def _norm_dims(dims):
    dims = tuple(int(d) for d in dims)
    if not 1 <= len(dims) <= 3 or min(dims) < 1:
        raise ValueError(f"bad dims {dims}")
    return dims


def _grid(*axes):
    """Fortran-order (first axis fastest) list of coordinates, shape (n, d)."""
    axes = [np.asarray(a, dtype=np.int64).reshape(-1) for a in axes]
    mesh = np.meshgrid(*axes, indexing="ij")
    return np.stack([m.reshape(-1, order="F") for m in mesh], axis=1)


class BetaTable:
    """Block table of the packing for one shape (see module docstring)."""

    def __init__(self, dims):
        self.dims = _norm_dims(dims)
        d = len(self.dims)
        info = [dim_info(n) for n in self.dims]
        self.M = [i[0] for i in info]
        self.S = [i[1] for i in info]
        self.P = [i[2] for i in info]
        R = [np.arange(1, p + 1) for p in self.P]
        H = [np.arange(m + 1, m + 1 + p) for m, p in zip(self.M, self.P)]
        coords, comps, verts = [], [], []

        def vertex(*c):
            coords.append(np.asarray([c], dtype=np.int64))
            comps.append(np.zeros(1, np.int8))
            verts.append(np.ones(1, bool))

        def pair(*axes):
            g = _grid(*axes)
            for comp in (0, 1):
                coords.append(g)
                comps.append(np.full(len(g), comp, np.int8))
                verts.append(np.zeros(len(g), bool))

        if d == 1:
            # mapping_cpu_kernels.jl:11-21
            for s in self.S[0]:
                vertex(s)
            pair(R[0])
        elif d == 2:
            # mapping_cpu_kernels.jl:38-87
            S1, S2 = self.S
            for s1 in S1:
                for s2 in S2:
                    vertex(s1, s2)
            for s1 in S1:
                pair([s1], R[1])
            for s2 in S2:
                pair(R[0], [s2])
            pair(R[0], R[1])
            pair(R[0], H[1])
        else:
            # mapping_cpu_kernels.jl:109-351
            S1, S2, S3 = self.S
            for s1 in S1:
                for s2 in S2:
                    for s3 in S3:
                        vertex(s1, s2, s3)
            for s1 in S1:          # edges along dim 3
                for s2 in S2:
                    pair([s1], [s2], R[2])
            for s1 in S1:          # edges along dim 2
                for s3 in S3:
                    pair([s1], R[1], [s3])
            for s2 in S2:          # edges along dim 1
                for s3 in S3:
                    pair(R[0], [s2], [s3])
            for s1 in S1:          # faces with a self-conjugate dim-1 index
                pair([s1], R[1], R[2])
                pair([s1], H[1], R[2])
            for s2 in S2:          # faces with a self-conjugate dim-2 index
                pair(R[0], [s2], R[2])
                pair(H[0], [s2], R[2])
            for s3 in S3:          # faces with a self-conjugate dim-3 index
                pair(R[0], R[1], [s3])
                pair(H[0], R[1], [s3])
            pair(R[0], R[1], R[2])  # interior
            pair(H[0], R[1], R[2])
            pair(R[0], H[1], R[2])
            pair(H[0], H[1], R[2])
        self.coord = np.concatenate(coords, axis=0)
        self.comp = np.concatenate(comps)
        self.vertex = np.concatenate(verts)
        self.N = int(np.prod(self.dims))
        assert len(self.coord) == self.N, (len(self.coord), self.N)
        # dense inverse: full-spectrum element -> beta index of its Re / Im part
        full = self.dims
        self.re_index = np.full(full, -1, np.int64)
        self.im_index = np.full(full, -1, np.int64)
        idx = np.arange(self.N)
        c = tuple(self.coord.T)
        re = self.comp == 0
        self.re_index[tuple(x[re] for x in c)] = idx[re]
        im = self.comp == 1
        self.im_index[tuple(x[im] for x in c)] = idx[im]
        # every represented element is hit exactly once per component
        assert (self.re_index >= 0).sum() == re.sum()
        assert (self.im_index >= 0).sum() == im.sum()

    def mirror(self, coord):
        coord = np.asarray(coord, dtype=np.int64)
        n = np.asarray(self.dims, dtype=np.int64)
        return (n - coord) % n

    @property
    def half_rows(self):
        return self.M[0] + 1


_TABLES: dict = {}


def table(dims) -> BetaTable:
    dims = _norm_dims(dims)
    if dims not in _TABLES:
        _TABLES[dims] = BetaTable(dims)
    return _TABLES[dims]


def spectrum_shape(dims, rdft=True):
    """Julia shape of the spectrum buffer: (M1+1, N2, N3) when rdft else dims."""
    dims = _norm_dims(dims)
    if rdft:
        return (dims[0] // 2 + 1,) + dims[1:]
    return dims


def dft_to_beta(v, dims):
    """beta = DFT_to_beta(v).  v: complex array in Julia (column-major) layout,
    either the half spectrum (M1+1, N2, N3) or the full spectrum (N1, N2, N3),
    given as a flat Fortran-ordered vector or an array of that shape (Fortran
    indexing).  Only rows 1..M1+1 of dim 1 are read (mapping_cpu_kernels.jl:109-351).
    """
    t = table(dims)
    v = np.asarray(v, dtype=np.complex128)
    rows = v.size // (t.N // t.dims[0])
    shape = (rows,) + t.dims[1:]
    vf = v.reshape(shape, order="F") if v.ndim == 1 else v
    coord = t.coord.copy()
    sign = np.ones(t.N)
    # represented elements with a second-half dim-1 index are read at their
    # conjugate partner, with the imaginary part negated (3D mirror blocks)
    far = coord[:, 0] > t.M[0]
    coord[far] = t.mirror(coord[far])
    sign[far & (t.comp == 1)] = -1.0
    vals = vf[tuple(coord.T)]
    part = np.where(t.comp == 0, vals.real, vals.imag)
    scale = np.where(t.vertex, 1.0, SQRT2) * sign   # +-sqrt(2) exactly
    out = np.where(t.vertex, part, scale * part)
    return out.astype(np.float64)


def beta_to_dft(beta, dims, rdft=True):
    """v = beta_to_DFT(beta); returns the flat Fortran-ordered complex vector of
    the half spectrum (rdft) or of the full spectrum (mapping_cpu_kernels.jl:353-738).
    """
    t = table(dims)
    beta = np.asarray(beta, dtype=np.float64).reshape(-1)
    if beta.size != t.N:
        raise ValueError("beta has wrong length")
    shape = spectrum_shape(t.dims, rdft)
    g = _grid(*[np.arange(n) for n in shape])          # Fortran order
    direct = t.re_index[tuple(g.T)] >= 0
    rep = np.where(direct[:, None], g, t.mirror(g))
    sigma = np.where(direct, 1.0, -1.0)
    ire = t.re_index[tuple(rep.T)]
    iim = t.im_index[tuple(rep.T)]
    assert (ire >= 0).all()
    vert = iim < 0
    re = np.where(vert, beta[ire], beta[ire] / SQRT2)
    im = np.zeros(len(g))
    nv = ~vert
    im[nv] = (sigma[nv] * beta[iim[nv]]) / SQRT2
    out = np.empty(len(g), np.complex128)
    out.real = re
    out.imag = im
    return out


def scatter_targets(dims):
    """Inverse view of dft_to_beta used by the fused GPU operator: for every
    half-spectrum element (flat Fortran index) the beta indices that read its
    real and imaginary parts (-1 when none) and the sign applied to Im."""
    t = table(dims)
    shape = spectrum_shape(t.dims, True)
    n = int(np.prod(shape))
    re_t = np.full(n, -1, np.int64)
    im_t = np.full(n, -1, np.int64)
    im_sign = np.ones(n)
    coord = t.coord.copy()
    far = coord[:, 0] > t.M[0]
    coord[far] = t.mirror(coord[far])
    flat = np.ravel_multi_index(tuple(coord.T), shape, order="F")
    idx = np.arange(t.N)
    r = t.comp == 0
    re_t[flat[r]] = idx[r]
    i = t.comp == 1
    im_t[flat[i]] = idx[i]
    im_sign[flat[i & far]] = -1.0
    return re_t, im_t, im_sign
