"""CPU FP64 oracle for the FFT operator M (TEST INFRASTRUCTURE ONLY).

Restates src/fft_utils.jl (rdft path, and with rdft=False the complex path):

* M_perp_beta  (:95-117)  y = mask .* (C2R(beta_to_DFT(beta)) / N * sqrt(N))
* M_perpt_z    (:73-93)   beta = DFT_to_beta(R2C(z) / sqrt(N))
* M_perpt_M_perp_vec (:119-123, :67-71)  M^T M beta
and the independent full-complex-FFT statement of test/fft_wei.jl:1-19.

Arrays are flat vectors in Julia column-major order (dimension 1 fastest).  A
flat Fortran vector of shape (N1, N2, N3) is the C-order array of shape
(N3, N2, N1), so the transforms below act on reversed shapes; the real-to-
complex halving then falls on the last C axis, i.e. on Julia's dimension 1.
FFTs come from scipy.fft (pocketfft, FP64).
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sfft

from . import maps
from . import maps_ref_forms as rf


def _cshape(dims):
    return tuple(int(d) for d in dims)[::-1]


def keep_mask(dims, missing):
    """1.0 at observed samples, 0.0 at `missing` (0-based linear indices)."""
    n = int(np.prod(dims))
    keep = np.ones(n)
    if missing is not None and len(missing):
        keep[np.asarray(missing, dtype=np.int64)] = 0.0
    return keep


def m_perp_beta(beta, dims, missing=None, workers=1, rdft=True):
    """rdft=False restates the complex path (fft_utils.jl:109-110): the full
    Hermitian spectrum, a complex inverse FFT, the real part."""
    n = int(np.prod(dims))
    if not rdft:
        v = maps.beta_to_dft(beta, dims, rdft=False)
        x = sfft.ifftn(v.reshape(_cshape(dims)), workers=workers).real.reshape(-1) * np.sqrt(n)
        if missing is not None and len(missing):
            x[np.asarray(missing, dtype=np.int64)] = 0.0
        return x
    v = maps.beta_to_dft(beta, dims, rdft=True)
    vc = v.reshape(_cshape(maps.spectrum_shape(dims, True)))
    x = sfft.irfftn(vc, s=_cshape(dims), workers=workers).reshape(-1)
    x = x * np.sqrt(n)
    if missing is not None and len(missing):
        x[np.asarray(missing, dtype=np.int64)] = 0.0
    return x


def m_perpt_z(z, dims, workers=1, rdft=True):
    """rdft=False: complex forward FFT of z (fft_utils.jl:80-81), the map
    reading the full spectrum."""
    n = int(np.prod(dims))
    zc = np.asarray(z, dtype=np.float64).reshape(_cshape(dims))
    if not rdft:
        v = sfft.fftn(zc.astype(np.complex128), workers=workers) / np.sqrt(n)
        return maps.dft_to_beta(v.reshape(-1), dims)
    v = sfft.rfftn(zc, workers=workers) / np.sqrt(n)
    return maps.dft_to_beta(v.reshape(-1), dims)


def mtm(beta, dims, missing=None, workers=1, rdft=True):
    return m_perpt_z(m_perp_beta(beta, dims, missing, workers, rdft), dims, workers, rdft)


# independent statement (test/fft_wei.jl:1-19): full complex FFTs + block maps
def m_perpt_z_wei(z, dims):
    n = int(np.prod(dims))
    zf = np.asarray(z, dtype=np.float64).reshape(dims, order="F")
    v = np.fft.fftn(zf) / np.sqrt(n)
    return rf.wei_dft_to_beta(v, tuple(dims))


def m_perp_beta_wei(beta, dims, missing=None):
    n = int(np.prod(dims))
    v = rf.wei_beta_to_dft(beta, tuple(dims))
    x = (np.fft.ifftn(v).real * np.sqrt(n)).reshape(-1, order="F")
    if missing is not None and len(missing):
        x[np.asarray(missing, dtype=np.int64)] = 0.0
    return x


def mtm_wei(beta, dims, missing=None):
    return m_perpt_z_wei(m_perp_beta_wei(beta, dims, missing), dims)


def dense_mtm(dims, missing=None):
    """Dense N x N matrix of M^T M (tiny shapes only): columns = mtm(e_i)."""
    n = int(np.prod(dims))
    out = np.empty((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        out[:, i] = mtm(e, dims, missing)
    return out
