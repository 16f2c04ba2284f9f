"""CompressedSensingADMM over libcsipm (src/admm.jl:8-61).

    CompressedSensingADMM(fft_parameter, fft_operator; rho=1, maxt=1000, tol=1e-6)

The whole iteration runs on the device (csipm_admm): the x-update CG on
(M^T M + rho I) through the fused operator passes, the soft threshold, the
dual update and the convergence norms in batched vector kernels; the host
reads one error per iteration.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib
from .fft_operator import FFTOperator, FFTParameters, _ptr


@dataclass
class ADMMStats:
    solution: torch.Tensor
    iterations: int
    cg_iterations: int
    status: str
    err: float
    seconds: float


def CompressedSensingADMM(fft_parameter: FFTParameters, fft_operator: FFTOperator, rho=1.0,
                          maxt=1000, tol=1e-6, stats=False):
    """Returns beta (= the reference's ztemp); with stats=True an ADMMStats."""
    op = fft_operator
    z0 = op._vec(fft_parameter.z0, dtype=torch.float64)
    beta = torch.empty(op.N, dtype=torch.float64, device=op.device)
    info = (ctypes.c_double * 8)()
    _lib.call("csipm_admm", op.handle, _ptr(z0), float(fft_parameter.lambda_), float(rho),
              int(maxt), float(tol), _ptr(beta), info)
    if not stats:
        return beta
    return ADMMStats(beta, int(info[1]), int(info[2]),
                     {0: "converged", 1: "maxt", 2: "failure"}[int(info[3])], info[4], info[6])
