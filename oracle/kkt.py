"""CPU FP64 oracle for the condensed FFT-KKT algebra and PCG (TEST INFRASTRUCTURE ONLY).

Restates src/fft_kkt.jl:
* CondensedFFTKKTSystem mul!  (:27-44)
* FFTPreconditioner mul!      (:73-85)
* build_kkt!                  (:285-307)
* solve_kkt! condensation / unpack (:309-354)
* full-KKT mul! (:213-251) and jtprod! (:253-275)
and the CG of Krylov.jl 0.10 called at :339 (x0 = 0, r = b, z = M r, p = z,
gamma = r'z; stop when sqrt(gamma) <= atol + rtol*sqrt(gamma0), on itmax, or on
non-positive curvature).
"""
from __future__ import annotations

import numpy as np


def kkt_mul(y, x, L1, L2, mtm_fn, alpha=1.0, beta=0.0):
    """y = alpha*K*x + beta*y, K = [[MtM + L1, L2], [L2, L1]] (fft_kkt.jl:27-44)."""
    n = len(L1)
    xb, xz = x[:n], x[n:]
    Mb = mtm_fn(xb)
    yb = beta * y[:n] + alpha * (Mb + L1 * xb + L2 * xz)
    yz = beta * y[n:] + alpha * (L2 * xb + L1 * xz)
    return np.concatenate([yb, yz])


def precond_build(sigma1, sigma2):
    """fft_kkt.jl:285-307: returns (L1, L2, P11, P12, P22)."""
    L1 = sigma1 + sigma2
    L2 = sigma1 - sigma2
    Minv = 1.0 / (1.0 + L1)
    S = 1.0 / (L1 - L2 ** 2 * Minv)
    P12 = -L2 * S * Minv
    P11 = Minv + Minv * L2 * S * L2 * Minv
    return L1, L2, P11, P12, S


def precond_mul(y, x, P11, P12, P22, alpha=1.0, beta=0.0):
    """fft_kkt.jl:73-85."""
    n = len(P11)
    xb, xz = x[:n], x[n:]
    yb = beta * y[:n] + alpha * (P11 * xb + P12 * xz)
    yz = beta * y[n:] + alpha * (P12 * xb + P22 * xz)
    return np.concatenate([yb, yz])


def condense_rhs(w, sigma1, sigma2):
    """fft_kkt.jl:333-335: w = [w1..w6] (6N) -> b (2N)."""
    n = len(sigma1)
    w1, w2, w3, w4, w5, w6 = (w[i * n:(i + 1) * n] for i in range(6))
    bb = w1 - w3 + w4 - sigma1 * w5 + sigma2 * w6
    bz = w2 - w3 - w4 - sigma1 * w5 - sigma2 * w6
    return np.concatenate([bb, bz])


def unpack_solution(w, x, sigma1, sigma2):
    """fft_kkt.jl:344-350: returns the solved 6N KKT vector."""
    n = len(sigma1)
    w = np.array(w, dtype=np.float64, copy=True)
    w1 = x[:n].copy()
    w2 = x[n:].copy()
    w3, w4, w5, w6 = (w[i * n:(i + 1) * n].copy() for i in range(2, 6))
    w5 = -(w3 + sigma1 * (w1 + w2 + w5))
    w6 = -(w4 + sigma2 * (-w1 + w2 + w6))
    w3 = (w3 + w5) / sigma1
    w4 = (w4 + w6) / sigma2
    return np.concatenate([w1, w2, w3, w4, w5, w6])


def full_kkt_mul(x, mtm_fn, n, alpha=1.0, beta=0.0, y=None):
    """fft_kkt.jl:239-246 (the part before MadNLP._kktmul!): x, y are 6N."""
    if y is None:
        y = np.zeros(6 * n)
    xb, xz, xs1, xs2, xy1, xy2 = (x[i * n:(i + 1) * n] for i in range(6))
    yb, yz, ys1, ys2, yy1, yy2 = (y[i * n:(i + 1) * n] for i in range(6))
    Mb = mtm_fn(xb)
    return np.concatenate([
        beta * yb + alpha * (Mb - xy1 + xy2),
        beta * yz - alpha * (xy1 + xy2),
        beta * ys1 - alpha * xy1,
        beta * ys2 - alpha * xy2,
        beta * yy1 + alpha * (-xb - xz - xs1),
        beta * yy2 + alpha * (xb - xz - xs2),
    ])


def jtprod(x, n):
    """fft_kkt.jl:253-275: J^T y for y = [y1; y2] (2N) -> 4N."""
    y1, y2 = x[:n], x[n:2 * n]
    return np.concatenate([-y1 + y2, -y1 - y2, -y1, -y2])


def pcg(A, b, M=None, atol=1e-12, rtol=0.0, itmax=None):
    """Preconditioned CG with Krylov.jl's cg conventions.  A, M: callables.
    Returns (x, iterations, sqrt(r'z), status) with status 'solved',
    'itmax' or 'zero-curvature'."""
    n = len(b)
    if itmax is None:
        itmax = 2 * n
    Mf = M if M is not None else (lambda r: r)
    x = np.zeros(n)
    r = np.array(b, dtype=np.float64, copy=True)
    z = Mf(r)
    p = z.copy()
    gamma = float(r @ z)
    rnorm = np.sqrt(max(gamma, 0.0))
    eps = atol + rtol * rnorm
    if rnorm <= eps:
        return x, 0, rnorm, "solved"
    it = 0
    while it < itmax:
        q = A(p)
        pq = float(p @ q)
        if pq <= 0.0:
            return x, it, rnorm, "zero-curvature"
        a = gamma / pq
        x = x + a * p
        r = r - a * q
        z = Mf(r)
        g_new = float(r @ z)
        it += 1
        rnorm = np.sqrt(max(g_new, 0.0))
        if rnorm <= eps:
            return x, it, rnorm, "solved"
        p = z + (g_new / gamma) * p
        gamma = g_new
    return x, it, rnorm, "itmax"
