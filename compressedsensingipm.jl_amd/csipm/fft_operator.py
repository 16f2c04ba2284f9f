"""FFTParameters / FFTOperator and the operator functions, over libcsipm.

Mirrors src/fft_utils.jl: the same names and argument meaning, with device
vectors held as torch tensors (torch is used only for device memory and the
current HIP stream; every computation runs in libcsipm's kernels).

    FFTParameters(DFTdim, DFTsize, z0, lambda_, index_missing)  fft_utils.jl:1-14
    FFTOperator(nbeta, DFTdim, DFTsize, index_missing, rdft)     fft_utils.jl:16-54
    M_perpt_z(op, z)            fft_utils.jl:73-93
    M_perp_beta(op, beta)       fft_utils.jl:95-117
    M_perpt_M_perp_vec(op, v)   fft_utils.jl:119-123
    DFT_to_beta(op, v) / beta_to_DFT(op, beta)   fft_utils.jl:144-222

`index_missing` holds 0-based linear indices (Julia's are 1-based).  Like the
reference, M_perp_beta and M_perpt_z return fresh results here; the
reference's buffer aliasing (fft_utils.jl:88-89) is internal scratch.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@dataclass
class FFTParameters:
    DFTdim: int
    DFTsize: tuple
    z0: object
    lambda_: float
    index_missing: np.ndarray

    def __post_init__(self):
        self.DFTsize = tuple(int(s) for s in self.DFTsize)
        if self.DFTdim != len(self.DFTsize):
            raise ValueError("DFTdim must equal len(DFTsize)")
        self.index_missing = np.asarray(self.index_missing, dtype=np.int64).reshape(-1)


class FFTOperator:
    """Device operator handle: twiddles, mask, scratch spectra, CG scalars.

    dtype float64 / float32 selects the operator arithmetic; mixed=True makes
    a handle whose operator and CG run in FP32 inside the FP64 interior-point
    driver (its operator entry points take float32 vectors)."""

    def __init__(self, nbeta, DFTdim, DFTsize, index_missing, rdft=True, dtype=torch.float64,
                 device="cuda", mixed=False):
        self.DFTdim = int(DFTdim)
        self.DFTsize = tuple(int(s) for s in DFTsize)
        self.N = int(np.prod(self.DFTsize))
        if int(nbeta) != self.N:
            raise ValueError("nbeta must equal prod(DFTsize)")
        # rdft=False (the reference's default, fft_utils.jl:46-48): the same
        # operator; the maps default to the full spectrum
        self.rdft = bool(rdft)
        if mixed:
            dtype = torch.float32     # the operator entry points run in FP32
        if dtype not in (torch.float64, torch.float32):
            raise ValueError("dtype must be float64 or float32")
        self.dtype = dtype
        self.mixed = bool(mixed)
        self.device = torch.device(device)
        prec = _lib.MIXED if mixed else (_lib.F64 if dtype == torch.float64 else _lib.F32)
        sizes = (ctypes.c_int64 * 3)(*(list(self.DFTsize) + [1] * (3 - len(self.DFTsize))))
        miss = np.ascontiguousarray(np.asarray(index_missing, dtype=np.int64).reshape(-1))
        self.index_missing = miss
        mp = miss.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)) if miss.size else None
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.call("csipm_op_create", ctypes.byref(h), self.DFTdim, sizes, int(self.rdft), prec, mp,
                      int(miss.size), _stream())
        self._h = h
        n = ctypes.c_int64()
        nc = ctypes.c_int64()
        p = ctypes.c_int()
        _lib.call("csipm_op_info", h, ctypes.byref(n), ctypes.byref(nc), ctypes.byref(p))
        self.Nc = int(nc.value)
        self.half_shape = (self.DFTsize[0] // 2 + 1,) + self.DFTsize[1:]

    @property
    def handle(self):
        self._sync_stream()
        return self._h

    def _sync_stream(self):
        _lib.call("csipm_op_set_stream", self._h, _stream())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().csipm_op_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # helpers -----------------------------------------------------------------
    def _vec(self, x, n=None, dtype=None):
        dtype = dtype or self.dtype
        n = self.N if n is None else n
        if not isinstance(x, torch.Tensor):
            x = torch.as_tensor(np.asarray(x), device=self.device)
        x = x.to(device=self.device, dtype=dtype).reshape(-1).contiguous()
        if x.numel() != n:
            raise ValueError(f"expected {n} entries, got {x.numel()}")
        return x

    def empty(self, n=None, dtype=None):
        return torch.empty(self.N if n is None else n, dtype=dtype or self.dtype,
                           device=self.device)


# ---------------------------------------------------------------------------
# operator functions (fft_utils.jl)
# ---------------------------------------------------------------------------

def M_perpt_z(op: FFTOperator, z):
    """beta = DFT_to_beta(R2C(z) / sqrt(N))   (fft_utils.jl:73-93)."""
    z = op._vec(z)
    out = op.empty()
    _lib.call("csipm_m_perpt_z", op.handle, _ptr(z), _ptr(out))
    return out


def M_perp_beta(op: FFTOperator, beta):
    """y = mask .* C2R(beta_to_DFT(beta)) / N * sqrt(N)   (fft_utils.jl:95-117)."""
    beta = op._vec(beta)
    out = op.empty()
    _lib.call("csipm_m_perp_beta", op.handle, _ptr(beta), _ptr(out))
    return out


def M_perpt_M_perp_vec(op: FFTOperator, vec):
    """M^T M vec, fused into one operator pass (fft_utils.jl:119-123)."""
    vec = op._vec(vec)
    out = op.empty()
    _lib.call("csipm_mtm", op.handle, _ptr(vec), _ptr(out))
    return out


def beta_to_DFT(op: FFTOperator, beta, full=None):
    """beta -> spectrum (complex128 tensor, Julia layout flattened);
    half spectrum (M1+1, N2, N3) for an rdft operator, else the full one
    (fft_utils.jl:195-222); `full` overrides."""
    full = (not op.rdft) if full is None else full
    beta = op._vec(beta, dtype=torch.float64)
    rows = op.DFTsize[0] if full else op.DFTsize[0] // 2 + 1
    n = rows * (op.N // op.DFTsize[0])
    out = torch.empty(2 * n, dtype=torch.float64, device=op.device)
    _lib.call("csipm_beta_to_dft", op.handle, _ptr(beta), _ptr(out), int(bool(full)))
    return torch.view_as_complex(out.view(n, 2))


def DFT_to_beta(op: FFTOperator, v, full=None):
    """spectrum -> beta (fft_utils.jl:144-174); v complex, Julia layout, the
    half spectrum for an rdft operator, else the full one; `full` overrides."""
    full = (not op.rdft) if full is None else full
    rows = op.DFTsize[0] if full else op.DFTsize[0] // 2 + 1
    n = rows * (op.N // op.DFTsize[0])
    if not isinstance(v, torch.Tensor):
        v = torch.as_tensor(np.asarray(v, dtype=np.complex128), device=op.device)
    v = v.to(device=op.device, dtype=torch.complex128).reshape(-1).contiguous()
    if v.numel() != n:
        raise ValueError(f"expected {n} spectrum entries, got {v.numel()}")
    vr = torch.view_as_real(v).contiguous()
    out = torch.empty(op.N, dtype=torch.float64, device=op.device)
    _lib.call("csipm_dft_to_beta", op.handle, _ptr(vr), _ptr(out), int(bool(full)))
    return out
