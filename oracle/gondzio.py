"""CPU FP64 oracle for the Gondzio formulation (TEST INFRASTRUCTURE ONLY).

Restates src/gondzio_model.jl and src/gondzio_kkt.jl.  Variables
x = [p; q; r] (N each), beta = p - q, p, q >= 0, r free; one equality block
c(x) = z0 - M (p - q) - r = 0; objective 1/2 r'r + lambda sum(p + q).
M = M_perp_beta (masked), M^T = M_perpt_z (as the reference: no mask).

* obj / grad / cons                       gondzio_model.jl:67-110
* jtprod      J^T v = [-M^T v; M^T v; -v]  gondzio_kkt.jl:243-262
* kkt_mul_full (the part of MadNLP.mul! the override adds)   :203-241
* kkt_mul (condensed K = [[M^T M + Wp, -M^T M], [-M^T M, M^T M + Wq]]) :28-53
* precond_build (build_kkt!: P = [[I + Wp, -I], [-I, I + Wq]]^-1)  :272-296
* precond_apply (GondzioPreconditioner mul!)                :80-92
* condense_rhs / unpack (solve_kkt!)                         :310-362
* pcg: Krylov.jl cg with atol = 1e-12, rtol = 0 (:339)
* solve: a primal-dual (Mehrotra) interior-point solve of the same
  problem on this Newton system (the reference delegates the loop to MadNLP).

Pinned by dense matrices on tiny shapes (explicit M), the congruence with
the condensed FFT system (A^T K A = [[M^T M + L1, L2], [L2, L1]] with
L1 = (Wp + Wq)/4, L2 = (Wp - Wq)/4, A = 1/2 [[I, I], [-I, I]]), and the
objective of the FFT-model oracle solve plus 1/2 ||z0||^2 (the two models
are the same LASSO when z0 is zero at the missing samples).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import kkt as fkkt
from . import operator as op


class GondzioOps:
    """M, M^T and M^T M of one shape and mask."""

    def __init__(self, dims, missing=None):
        self.dims = tuple(int(d) for d in dims)
        self.N = int(np.prod(self.dims))
        self.missing = np.zeros(0, np.int64) if missing is None else np.asarray(missing)

    def M(self, v):
        return op.m_perp_beta(v, self.dims, self.missing)

    def Mt(self, z):
        return op.m_perpt_z(z, self.dims)

    def MtM(self, v):
        return op.mtm(v, self.dims, self.missing)


def obj(x, lam, N):
    r = x[2 * N:]
    return 0.5 * float(r @ r) + lam * float(np.sum(x[:2 * N]))


def grad(x, lam, N):
    g = np.empty_like(x)
    g[:2 * N] = lam
    g[2 * N:] = x[2 * N:]
    return g


def cons(o: GondzioOps, x, z0):
    N = o.N
    return z0 - o.M(x[:N] - x[N:2 * N]) - x[2 * N:]


def jtprod(o: GondzioOps, v):
    t = o.Mt(v)
    return np.concatenate([-t, t, -v])


def kkt_mul_full(o: GondzioOps, x, y, alpha, beta):
    """y (4N: [p; q; r; y]) = alpha*A x + beta*y, A without MadNLP's diagonals."""
    N = o.N
    xp, xq, xr, xy = x[:N], x[N:2 * N], x[2 * N:3 * N], x[3 * N:]
    t = o.Mt(xy)
    s = o.M(xq - xp)
    out = np.concatenate([-alpha * t, alpha * t, alpha * (xr - xy), alpha * (s - xr)])
    return out + beta * y


def kkt_mul(o: GondzioOps, x, y, Wp, Wq, alpha, beta):
    N = o.N
    p, q = x[:N], x[N:]
    m = o.MtM(p - q)
    return np.concatenate([alpha * (Wp * p + m) + beta * y[:N],
                           alpha * (Wq * q - m) + beta * y[N:]])


def precond_build(reg_p, reg_q):
    inv = 1.0 / (1.0 + reg_p)
    S = 1.0 / (1.0 + reg_q - inv)
    P12 = inv * S
    P11 = inv + inv * S * inv
    return P11, P12, S


def precond_apply(x, y, P11, P12, P22, alpha, beta):
    N = len(P11)
    xp, xq = x[:N], x[N:]
    return np.concatenate([alpha * (P11 * xp + P12 * xq) + beta * y[:N],
                           alpha * (P12 * xp + P22 * xq) + beta * y[N:]])


def condense_rhs(o: GondzioOps, w):
    """w = [w_p; w_q; w_r; w_y] -> rhs (2N)."""
    N = o.N
    t = o.Mt(w[2 * N:3 * N] + w[3 * N:])
    return np.concatenate([w[:N] - t, w[N:2 * N] + t])


def unpack(o: GondzioOps, w, dx):
    """w_x = dx; dr = -w_y - M (dp - dq); dy = dr - w_r (in place, returned)."""
    N = o.N
    w = w.copy()
    wr = w[2 * N:3 * N].copy()
    w[:2 * N] = dx
    dr = -w[3 * N:] - o.M(dx[:N] - dx[N:])
    w[2 * N:3 * N] = dr
    w[3 * N:] = dr - wr
    return w


def pcg(o: GondzioOps, b, Wp, Wq, P, atol=1e-12, rtol=0.0, itmax=0):
    """Krylov.jl cg on the condensed K with preconditioner P = (P11, P12, P22)."""
    K = lambda v: kkt_mul(o, v, np.zeros_like(v), Wp, Wq, 1.0, 0.0)   # noqa: E731
    M = lambda v: precond_apply(v, np.zeros_like(v), *P, 1.0, 0.0)    # noqa: E731
    return fkkt.pcg(K, b, M, atol=atol, rtol=rtol, itmax=itmax or 2 * len(b))


@dataclass
class GondzioResult:
    x: np.ndarray          # [p; q; r]
    y: np.ndarray          # multiplier of c(x) = 0
    objective: float
    iterations: int
    status: str


def solve(o: GondzioOps, z0, lam, tol=1e-10, max_iter=200, cg_atol=1e-13):
    """Mehrotra predictor-corrector on p, q >= 0 with r = z0 - M (p - q)
    eliminated; every Newton system is the condensed Gondzio K, solved by
    `pcg` with build_kkt!'s preconditioner (Wp = zp/p, Wq = zq/q)."""
    N = o.N
    b = o.Mt(z0)
    p = np.ones(N)
    q = np.ones(N)
    zp = np.full(N, max(lam, 1.0))
    zq = np.full(N, max(lam, 1.0))
    status = "max_iter"
    it = 0
    for it in range(1, max_iter + 1):
        m = o.MtM(p - q) - b
        gp, gq = m + lam - zp, -m + lam - zq
        mu = (p @ zp + q @ zq) / (2 * N)
        scale = max(1.0, np.abs(b).max(), lam)
        if max(np.abs(gp).max(), np.abs(gq).max()) / scale <= tol and mu <= tol * scale:
            status = "solved"
            break
        Wp, Wq = zp / p, zq / q
        P = precond_build(Wp, Wq)

        def direction(sp, sq):
            rhs = np.concatenate([-gp - sp / p, -gq - sq / q])
            dx = pcg(o, rhs, Wp, Wq, P, atol=cg_atol * max(1.0, np.abs(rhs).max()))[0]
            dp, dq = dx[:N], dx[N:]
            return dp, dq, -(sp + zp * dp) / p, -(sq + zq * dq) / q

        def step(v, dv):
            neg = dv < 0
            return min(1.0, float(np.min(-v[neg] / dv[neg]))) if neg.any() else 1.0

        dp, dq, dzp, dzq = direction(p * zp, q * zq)
        ap = min(step(p, dp), step(q, dq))
        ad = min(step(zp, dzp), step(zq, dzq))
        mu_aff = ((p + ap * dp) @ (zp + ad * dzp) + (q + ap * dq) @ (zq + ad * dzq)) / (2 * N)
        sigma = (mu_aff / mu) ** 3
        dp, dq, dzp, dzq = direction(p * zp + dp * dzp - sigma * mu,
                                     q * zq + dq * dzq - sigma * mu)
        ap = 0.99 * min(step(p, dp), step(q, dq))
        ad = 0.99 * min(step(zp, dzp), step(zq, dzq))
        p, q = p + ap * dp, q + ap * dq
        zp, zq = zp + ad * dzp, zq + ad * dzq
    r = z0 - o.M(p - q)
    x = np.concatenate([p, q, r])
    return GondzioResult(x, r.copy(), obj(x, lam, N), it, status)
