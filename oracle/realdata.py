"""CPU oracle for the real-data front end (TEST INFRASTRUCTURE ONLY).

Restates, in plain Python / numpy:

* index_missing of benchmarks/mastodonte.jl:38-50: loops k, j, i over the
  mask (i fastest) collecting i + (j-1) N1 + (k-1) N1 N2 where mask == 0
  (1-based there; 0-based here) -- `index_missing_loops`, and its
  vectorised equivalent `index_missing`;
* punch_3D_cart of benchmarks/crystal.jl:13-27: the grid points inside an
  ellipsoid, by the same filter over every grid point -- `punch_loops`;
* the solution dump of mastodonte.jl:109-116,
  x = sqrt(N) real(ifft(beta_to_DFT(beta))) -- `reconstruct`, through the
  full-spectrum map and numpy's complex inverse FFT.

Pinning: the loops against the vectorised forms on small inputs;
`reconstruct` against the maskless oracle operator (a second route: the
real-to-complex C2R of the half spectrum) and the round trip
reconstruct(DFT_to_beta(fft(x)/sqrt(N))) = x for real x, the definition of
beta_true in test/ipm_example_3D.jl:82-83.
"""
from __future__ import annotations

import itertools

import numpy as np

from . import maps


def index_missing_loops(mask3):
    """mastodonte.jl:38-50 literally (0-based); mask3 indexed [i, j, k]."""
    n1, n2, n3 = mask3.shape
    out = []
    for k in range(n3):
        for j in range(n2):
            for i in range(n1):
                if mask3[i, j, k] == 0:
                    out.append(i + j * n1 + k * n1 * n2)
    return np.asarray(out, dtype=np.int64)


def index_missing(mask_colmajor):
    """Zero entries of a column-major flattened mask, 0-based, increasing."""
    return np.flatnonzero(np.asarray(mask_colmajor).reshape(-1) == 0).astype(np.int64)


def punch_loops(center, radius, x, y, z):
    """crystal.jl:13-27: filter over CartesianIndices((1:nx, 1:ny, 1:nz)),
    returned as 0-based column-major linear indices in that iteration order
    (first index fastest)."""
    rx, ry, rz = radius if isinstance(radius, (tuple, list)) else (radius,) * 3
    nx, ny, nz = len(x), len(y), len(z)
    out = []
    for k, j, i in itertools.product(range(nz), range(ny), range(nx)):
        if (((x[i] - center[0]) / rx) ** 2 + ((y[j] - center[1]) / ry) ** 2
                + ((z[k] - center[2]) / rz) ** 2) <= 1.0:
            out.append(i + nx * (j + ny * k))
    return np.asarray(out, dtype=np.int64)


def reconstruct(beta, dims):
    """sqrt(N) * real(ifft(beta_to_DFT(beta))) over the full spectrum."""
    dims = tuple(int(d) for d in dims)
    n = int(np.prod(dims))
    v = maps.beta_to_dft(np.asarray(beta, dtype=np.float64), dims, rdft=False)
    v = np.asarray(v).reshape(-1)
    arr = v.reshape(tuple(reversed(dims)))          # Julia column-major view
    x = np.fft.ifftn(arr) * np.sqrt(n)
    return np.real(x).reshape(-1)


def beta_of_signal(x, dims):
    """beta_true = DFT_to_beta(fft(x) / sqrt(N)) (test/ipm_example_3D.jl:82-83)."""
    dims = tuple(int(d) for d in dims)
    n = int(np.prod(dims))
    arr = np.asarray(x, dtype=np.float64).reshape(tuple(reversed(dims)))
    return maps.dft_to_beta((np.fft.fftn(arr) / np.sqrt(n)).reshape(-1), dims)
