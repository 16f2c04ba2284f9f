"""CPU FP64 oracle for the full solve (TEST INFRASTRUCTURE ONLY).

The reference solves FFTNLPModel (src/fft_model.jl:66-101)

    min  f(beta, c) = 1/2 ||M beta||^2 - beta' (M^T z0) + lambda * sum(c)
    s.t. -beta - c <= 0,   beta - c <= 0

with MadNLP's interior-point method, every Newton system being reduced to the
condensed FFT-KKT system of src/fft_kkt.jl (K = [[MtM + L1, L2], [L2, L1]],
L1 = S1 + S2, L2 = S1 - S2, S_i the slack barrier diagonals, :285-307) and
solved by preconditioned CG (:339).  MadNLP is not vendored and cannot run
here, so its iterate sequence is not reproduced; this module is the project's
own primal-dual method on the *same* KKT system, written as the specification
the GPU driver follows step for step:

state   beta, c (slacks t1 = c + beta > 0, t2 = c - beta > 0), duals z1, z2 > 0
resid   r_b = MtM beta - b - z1 + z2,   r_c = lambda - z1 - z2,
        mu = (t1'z1 + t2'z2) / (2N)
step    Mehrotra predictor-corrector.  For targets g_i = (s*mu - t_i z_i - corr_i)/t_i
        solve K [db; dc] = [-r_b + g1 - g2; -r_c + g1 + g2] by PCG with the
        reference preconditioner, then dt1 = dc + db, dt2 = dc - db,
        dz_i = g_i - S_i dt_i.  Predictor: s = 0, corr = 0; sigma = min(1, (mu_aff/mu)^3);
        corrector: s = sigma, corr_i = dt_i^aff dz_i^aff.
length  fraction to the boundary tau = 0.995 on the primal (t) and dual (z)
        sides separately.
stop    max(||r_b||_inf, ||r_c||_inf) <= tol * (1 + ||b||_inf) and mu <= tol * (1 + |f|/N)
        (see `converged`), or max_iter.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import kkt


@dataclass
class IPMOptions:
    tol: float = 1e-8
    max_iter: int = 200
    cg_atol: float = 0.0
    cg_rtol: float = 1e-10
    cg_itmax: int = 1000
    tau: float = 0.995


@dataclass
class IPMResult:
    beta: np.ndarray
    c: np.ndarray
    z1: np.ndarray
    z2: np.ndarray
    objective: float
    iterations: int
    status: str
    cg_iterations: list = field(default_factory=list)
    history: list = field(default_factory=list)


def objective(beta, c, Qbeta, b, lam):
    """fft_model.jl:73: 1/2 ||M beta||^2 - beta'M^T z0 + lambda sum(c), using
    ||M beta||^2 = beta' M^T M beta."""
    return 0.5 * float(beta @ Qbeta) - float(beta @ b) + lam * float(np.sum(c))


def initial_point(b, lam):
    """beta = 0, c = 1 + ||b||_inf, z1 = z2 = max(lambda/2, ...) so that
    r_c = 0 when lambda > 0."""
    n = len(b)
    scale = 1.0 + float(np.max(np.abs(b))) if n else 1.0
    beta = np.zeros(n)
    c = np.full(n, scale)
    zval = lam / 2.0 if lam > 0 else 1.0
    z1 = np.full(n, zval)
    z2 = np.full(n, zval)
    return beta, c, z1, z2


def max_step(x, dx):
    """Largest a in (0, 1] with x + a*dx >= 0 (x > 0)."""
    neg = dx < 0
    if not np.any(neg):
        return 1.0
    return float(min(1.0, np.min(-x[neg] / dx[neg])))


def converged(rb, rc, mu, f, bnorm, n, tol):
    dual = max(float(np.max(np.abs(rb))), float(np.max(np.abs(rc))))
    return dual <= tol * (1.0 + bnorm) and mu <= tol * (1.0 + abs(f) / n)


def solve(b, lam, mtm_fn, opts: IPMOptions | None = None, verbose=False):
    """Solve the compressed-sensing QP for b = M^T z0 with operator mtm_fn."""
    opts = opts or IPMOptions()
    n = len(b)
    bnorm = float(np.max(np.abs(b))) if n else 0.0
    beta, c, z1, z2 = initial_point(b, lam)
    cg_its = []
    hist = []
    status = "max_iter"
    it = 0
    f = 0.0
    for it in range(opts.max_iter + 1):
        Qb = mtm_fn(beta)
        t1 = c + beta
        t2 = c - beta
        rb = Qb - b - z1 + z2
        rc = lam - z1 - z2
        mu = (float(t1 @ z1) + float(t2 @ z2)) / (2 * n)
        f = objective(beta, c, Qb, b, lam)
        hist.append((it, f, float(np.max(np.abs(rb))), float(np.max(np.abs(rc))), mu))
        if verbose:
            print(f"it {it:3d} f {f:+.12e} rb {hist[-1][2]:.2e} rc {hist[-1][3]:.2e} mu {mu:.2e}")
        if converged(rb, rc, mu, f, bnorm, n, opts.tol):
            status = "solved"
            break
        if it == opts.max_iter:
            break
        s1 = z1 / t1
        s2 = z2 / t2
        L1, L2, P11, P12, P22 = kkt.precond_build(s1, s2)

        def K(x):
            return kkt.kkt_mul(np.zeros(2 * n), x, L1, L2, mtm_fn)

        def P(x):
            return kkt.precond_mul(np.zeros(2 * n), x, P11, P12, P22)

        def direction(g1, g2):
            rhs = np.concatenate([-rb + g1 - g2, -rc + g1 + g2])
            d, k, _, _ = kkt.pcg(K, rhs, P, atol=opts.cg_atol, rtol=opts.cg_rtol,
                                 itmax=opts.cg_itmax)
            cg_its.append(k)
            db, dc = d[:n], d[n:]
            dt1 = dc + db
            dt2 = dc - db
            dz1 = g1 - s1 * dt1
            dz2 = g2 - s2 * dt2
            return db, dc, dt1, dt2, dz1, dz2

        # predictor
        db, dc, dt1, dt2, dz1, dz2 = direction(-z1, -z2)
        ap = min(max_step(t1, dt1), max_step(t2, dt2))
        ad = min(max_step(z1, dz1), max_step(z2, dz2))
        mu_aff = (float((t1 + ap * dt1) @ (z1 + ad * dz1)) +
                  float((t2 + ap * dt2) @ (z2 + ad * dz2))) / (2 * n)
        sigma = min(1.0, (mu_aff / mu) ** 3)
        # corrector
        g1 = (sigma * mu - t1 * z1 - dt1 * dz1) / t1
        g2 = (sigma * mu - t2 * z2 - dt2 * dz2) / t2
        db, dc, dt1, dt2, dz1, dz2 = direction(g1, g2)
        tau = opts.tau
        ap = min(1.0, tau * min(max_step(t1, dt1), max_step(t2, dt2)))
        ad = min(1.0, tau * min(max_step(z1, dz1), max_step(z2, dz2)))
        beta = beta + ap * db
        c = c + ap * dc
        z1 = z1 + ad * dz1
        z2 = z2 + ad * dz2
    return IPMResult(beta, c, z1, z2, f, it, status, cg_its, hist)


# --------------------------------------------------------------------------
# algorithm-free LASSO optimality certificate
# --------------------------------------------------------------------------

def soft_threshold(x, t):
    return np.sign(x) * np.maximum(np.abs(x) - t, 0.0)


def lasso_certificate(beta, b, lam, mtm_fn, m_perp_fn, z0):
    """Checks beta against  P(beta) = 1/2||M beta - z0||^2 + lam ||beta||_1
    (= f(beta, |beta|) + 1/2||z0||^2 when z0 vanishes at missing samples):
    returns (prox-gradient fixed-point residual, duality gap, P(beta)).

    The prox step uses step 1, valid because ||M^T M|| <= 1.  The dual point
    is the scaled residual u = (z0 - M beta) * min(1, lam / ||M^T(z0 - M beta)||_inf)
    with D(u) = 1/2||z0||^2 - 1/2||z0 - u||^2, so 0 <= P(beta) - P* <= gap."""
    Qb = mtm_fn(beta)
    grad = Qb - b
    fp = float(np.max(np.abs(beta - soft_threshold(beta - grad, lam))))
    Mb = m_perp_fn(beta)
    res = z0 - Mb
    primal = 0.5 * float(res @ res) + lam * float(np.sum(np.abs(beta)))
    Atr = b - Qb                      # M^T (z0 - M beta)
    amax = float(np.max(np.abs(Atr)))
    s = 1.0 if amax <= lam or amax == 0.0 else lam / amax
    u = s * res
    dual = 0.5 * float(z0 @ z0) - 0.5 * float((z0 - u) @ (z0 - u))
    return fp, primal - dual, primal
