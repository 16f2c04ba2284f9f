"""Real-data front end: masks, punched lattices, array files, reconstruction.

The reference's real-data drivers (benchmarks/mastodonte.jl, crystal.jl) read
an observed 3D signal and a sampling mask, turn the mask into
`index_missing`, solve, and write the recovered signal
x = sqrt(N) * real(ifft(beta_to_DFT(beta))).  This module does the same
around libcsipm:

* index_missing_from_mask: the zero entries of a mask / weights array as
  0-based linear indices, on the GPU (csipm_mask_index_missing), in the order
  of mastodonte.jl:38-50's loops;
* punch_3D_cart / punch_lattice: the ellipsoidal punches of crystal.jl:13-27
  and :30-80 on a coordinate grid (host; data preparation);
* reconstruct: csipm_reconstruct (M_perp_beta without the mask);
* load_array / save_array: .npy / .npz (HDF5 is not in this environment);
  arrays are converted to Julia's column-major order;
* mastodonte: mask + signal -> solve -> recovered signal.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .fft_operator import FFTOperator, FFTParameters, _ptr, _stream


def index_missing_from_mask(mask, device="cuda") -> torch.Tensor:
    """0-based linear (column-major) indices of the zero entries of `mask`.

    mask: a torch tensor or numpy array already in Julia's column-major
    element order (flattened), or any array with `.ravel(order="F")`
    semantics applied by the caller.  Returns an int64 tensor on `device`
    (mastodonte.jl:38-50 builds the same list, 1-based, with three loops)."""
    if not isinstance(mask, torch.Tensor):
        mask = torch.as_tensor(np.asarray(mask, dtype=np.float64))
    w = mask.to(device=device, dtype=torch.float64).reshape(-1).contiguous()
    n = w.numel()
    out = torch.empty(max(n, 1), dtype=torch.int64, device=w.device)
    cnt = ctypes.c_int64()
    with torch.cuda.device(w.device):
        _lib.call("csipm_mask_index_missing", _ptr(w), int(n), _ptr(out), ctypes.byref(cnt),
                  _stream())
    return out[:cnt.value]


def to_column_major(a: np.ndarray) -> np.ndarray:
    """Flatten an array indexed a[i1, i2, i3] (Julia's (N1, N2, N3) order)
    into Julia's memory order (first index fastest)."""
    return np.asarray(a).ravel(order="F")


def from_column_major(x, dims) -> np.ndarray:
    """Inverse of to_column_major: vector -> array indexed a[i1, i2, i3]."""
    return np.asarray(x).reshape(tuple(dims), order="F")


def punch_3D_cart(center, radius, x, y, z, linear=True):
    """Grid points (x[i], y[j], z[k]) inside the ellipsoid of `center` and
    `radius` (a number or a 3-tuple): crystal.jl:13-27.  Returns 0-based
    column-major linear indices (linear=True) or an (m, 3) index array."""
    rx, ry, rz = radius if isinstance(radius, (tuple, list)) else (radius,) * 3
    x, y, z = (np.asarray(v, dtype=np.float64) for v in (x, y, z))
    gx = ((x - center[0]) / rx) ** 2
    gy = ((y - center[1]) / ry) ** 2
    gz = ((z - center[2]) / rz) ** 2
    inside = gx[:, None, None] + gy[None, :, None] + gz[None, None, :] <= 1.0
    if not inside.any():
        raise ValueError("Empty punch.")
    if not linear:
        return np.argwhere(inside)
    return np.flatnonzero(inside.ravel(order="F"))


def punch_lattice(x, y, z, centers, radius):
    """Union of the punches of all `centers` (crystal.jl:30-80 punches every
    integer lattice point of the box); sorted unique 0-based indices."""
    idx = [punch_3D_cart(c, radius, x, y, z) for c in centers]
    return np.unique(np.concatenate(idx)) if idx else np.zeros(0, np.int64)


def load_array(path, key=None) -> np.ndarray:
    """A .npy / .npz array (no pickled objects)."""
    if str(path).endswith(".npz"):
        with np.load(path, allow_pickle=False) as f:
            return np.asarray(f[key or list(f.keys())[0]])
    return np.load(path, allow_pickle=False)


def save_array(path, a):
    np.save(path, np.asarray(a), allow_pickle=False)


def reconstruct(op: FFTOperator, beta) -> torch.Tensor:
    """x = sqrt(N) real(ifft(beta_to_DFT(beta))), all samples kept
    (mastodonte.jl:109-116)."""
    beta = op._vec(beta)
    out = op.empty()
    _lib.call("csipm_reconstruct", op.handle, _ptr(beta), _ptr(out))
    return out


@dataclass
class RealDataResult:
    beta: torch.Tensor
    recovered: torch.Tensor
    objective: float
    iterations: int
    cg_iterations: int
    status: str
    seconds: float
    index_missing: np.ndarray


def mastodonte(z0, mask, lambda_=1e-6, rdft=False, max_iter=23, tol=1e-8, mixed=False,
               device="cuda") -> RealDataResult:
    """The reference's real-data driver (mastodonte.jl:11-71, :98-117): mask
    -> index_missing, solve the FFT model (max_iter 23, tol 1e-8 as there),
    recover x = sqrt(N) real(ifft(beta_to_DFT(beta))).  z0 and mask are
    arrays indexed [i1, i2, i3] (Julia's order) or already flattened
    column-major with `mask.shape` giving the dims."""
    from .model import FFTNLPModel, ipm_solve
    mask = np.asarray(mask)
    dims = tuple(int(s) for s in mask.shape)
    w = to_column_major(mask).astype(np.float64)
    miss = index_missing_from_mask(w, device=device).cpu().numpy()
    z = to_column_major(np.asarray(z0, dtype=np.float64))
    params = FFTParameters(len(dims), dims, torch.from_numpy(z).to(device), lambda_, miss)
    t0 = time.perf_counter()
    nlp = FFTNLPModel(params, rdft=rdft, mixed=mixed)
    st = ipm_solve(nlp, tol=tol, max_iter=max_iter)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    op = FFTOperator(int(np.prod(dims)), len(dims), dims, np.zeros(0, np.int64), rdft=rdft)
    x = reconstruct(op, st.solution.to(torch.float64))
    return RealDataResult(st.solution, x, st.objective, st.iterations, st.cg_iterations,
                          st.status, secs, miss)
