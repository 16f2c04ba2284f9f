"""CPU FP64 oracle for the ADMM solver (TEST INFRASTRUCTURE ONLY).

Restates CompressedSensingADMM (src/admm.jl:8-61) for

    min 1/2 ||M beta - z0||^2 + lambda ||beta||_1

x-update: CG on (M^T M + rho I) x = M^T z0 + rho z - y  (admm.jl:1-6, 29-36),
with Krylov.jl's cg defaults (x from 0, stop at ||r|| <= atol + rtol ||b||,
atol = rtol = sqrt(eps), at most 2n iterations);
z = softthreshold(x + y/rho, lambda/rho) (:40-41, :64-77); y += rho (x - z)
(:44); err = max(||x1 - x0||, ||y1 - y0||, ||z1 - z0||) (:47-53); loop while
t < maxt and err > tol (:32); x0 = z0 = ones, y0 = 0 (:19-24); returns z.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import kkt


@dataclass
class ADMMResult:
    beta: np.ndarray
    iterations: int
    cg_iterations: int
    err: float
    converged: bool


def soft_threshold(x, t):
    """admm.jl:64-77: x - t above t, x + t below -t, 0 between."""
    return np.where(x > t, x - t, np.where(x < -t, x + t, 0.0))


def solve(b0, lam, mtm_fn, rho=1.0, maxt=1000, tol=1e-6):
    """b0 = M^T z0; mtm_fn(v) = M^T M v."""
    n = len(b0)
    eps = np.sqrt(np.finfo(np.float64).eps)
    x0 = np.ones(n)
    z = np.ones(n)
    y = np.zeros(n)
    t = 0
    err = 1.0
    cg_total = 0
    ident = lambda r: r                                  # noqa: E731
    op = lambda v: mtm_fn(v) + rho * v                   # noqa: E731
    while t < maxt and err > tol:
        b = b0 + rho * z - y
        x1, its, _, _ = kkt.pcg(op, b, ident, atol=eps, rtol=eps, itmax=2 * n)
        cg_total += its
        z1 = soft_threshold(x1 + y / rho, lam / rho)
        y1 = y + rho * (x1 - z1)
        err = max(np.linalg.norm(x1 - x0), np.linalg.norm(y1 - y), np.linalg.norm(z1 - z))
        x0, z, y = x1, z1, y1
        t += 1
    return ADMMResult(z, t, cg_total, float(err), err <= tol)
