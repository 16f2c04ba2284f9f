"""CPU checks that pin the Gondzio oracle (oracle/gondzio.py).

* kkt_mul, kkt_mul_full, jtprod, cons against explicit dense matrices of M
  (masked M_perp_beta) and M^T (M_perpt_z) on tiny shapes.
* precond_build: P times [[I + Wp, -I], [-I, I + Wq]] is the identity.
* condense_rhs + pcg + unpack solve the reduced 4N KKT system of
  gondzio_kkt.jl:316-327 (checked against a dense solve).
* The congruence the library's Gondzio CG relies on: A^T K A equals the
  condensed FFT system with L1 = (Wp + Wq)/4, L2 = (Wp - Wq)/4.
* The oracle's Gondzio interior-point solve reaches the FFT-model oracle's
  objective plus 1/2 ||z0||^2 (the same LASSO); a planted mask-free M^T M
  moves it.
"""
import numpy as np
import pytest

from oracle import gondzio as gz
from oracle import ipm, kkt
from oracle import operator as op


def _dense(fn, n):
    out = np.empty((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        out[:, i] = fn(e)
    return out


SHAPES = [(12,), (11,), (6, 4), (5, 3), (4, 3, 2), (3, 3, 3)]


def _setup(dims, seed=0):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    miss = np.sort(rng.choice(n, size=max(1, n // 5), replace=False))
    o = gz.GondzioOps(dims, miss)
    M = _dense(o.M, n)
    Mt = _dense(o.Mt, n)
    return rng, n, o, M, Mt


@pytest.mark.parametrize("dims", SHAPES)
def test_products_dense(dims):
    rng, n, o, M, Mt = _setup(dims)
    Wp, Wq = rng.random(n) + 0.1, rng.random(n) + 0.1
    x = rng.standard_normal(2 * n)
    y = rng.standard_normal(2 * n)
    K = np.block([[Mt @ M + np.diag(Wp), -Mt @ M], [-Mt @ M, Mt @ M + np.diag(Wq)]])
    assert np.allclose(gz.kkt_mul(o, x, y, Wp, Wq, 0.7, -0.3), 0.7 * K @ x - 0.3 * y,
                       atol=1e-12, rtol=0)
    v = rng.standard_normal(n)
    assert np.allclose(gz.jtprod(o, v), np.concatenate([-Mt @ v, Mt @ v, -v]), atol=1e-12)
    z0 = rng.standard_normal(n)
    x3 = rng.standard_normal(3 * n)
    assert np.allclose(gz.cons(o, x3, z0), z0 - M @ (x3[:n] - x3[n:2 * n]) - x3[2 * n:],
                       atol=1e-12)
    x4 = rng.standard_normal(4 * n)
    y4 = rng.standard_normal(4 * n)
    Z = np.zeros((n, n))
    I = np.eye(n)
    A = np.block([[Z, Z, Z, -Mt], [Z, Z, Z, Mt], [Z, Z, Z, -I], [-M, M, -I, Z]])
    A[2 * n:3 * n, 2 * n:3 * n] = I      # y_r = alpha (x_r - x_y)
    assert np.allclose(gz.kkt_mul_full(o, x4, y4, 1.3, 0.5), 1.3 * A @ x4 + 0.5 * y4,
                       atol=1e-12)


def test_precond_inverse():
    rng = np.random.default_rng(2)
    n = 50
    Wp, Wq = 10.0 ** rng.uniform(-8, 8, n), 10.0 ** rng.uniform(-8, 8, n)
    P11, P12, P22 = gz.precond_build(Wp, Wq)
    for i in range(n):
        P = np.array([[P11[i], P12[i]], [P12[i], P22[i]]])
        B = np.array([[1 + Wp[i], -1.0], [-1.0, 1 + Wq[i]]])
        assert np.allclose(P @ B, np.eye(2), atol=1e-6)


@pytest.mark.parametrize("dims", SHAPES)
def test_solve_kkt_dense(dims):
    rng, n, o, M, Mt = _setup(dims, 1)
    Wp, Wq = rng.random(n) + 0.05, rng.random(n) + 0.05
    w = rng.standard_normal(4 * n)
    rhs = gz.condense_rhs(o, w)
    dx, it, res, st = gz.pcg(o, rhs, Wp, Wq, gz.precond_build(Wp, Wq), atol=1e-14)
    assert st == "solved"
    sol = gz.unpack(o, w, dx)
    Z = np.zeros((n, n))
    I = np.eye(n)
    A = np.block([[np.diag(Wp), Z, Z, -Mt], [Z, np.diag(Wq), Z, Mt], [Z, Z, I, -I],
                  [-M, M, -I, Z]])
    ref = np.linalg.solve(A, w)
    assert np.allclose(sol, ref, atol=1e-9 * np.abs(ref).max())


@pytest.mark.parametrize("dims", SHAPES)
def test_congruence_with_fft_system(dims):
    rng, n, o, M, Mt = _setup(dims, 3)
    Wp, Wq = rng.random(n) + 0.1, rng.random(n) + 0.1
    yv = rng.standard_normal(2 * n)
    x = 0.5 * np.concatenate([yv[:n] + yv[n:], yv[n:] - yv[:n]])          # A y
    Kx = gz.kkt_mul(o, x, np.zeros(2 * n), Wp, Wq, 1.0, 0.0)
    AtKx = 0.5 * np.concatenate([Kx[:n] - Kx[n:], Kx[:n] + Kx[n:]])      # A^T K A y
    L1, L2 = (Wp + Wq) / 4, (Wp - Wq) / 4
    ref = kkt.kkt_mul(np.zeros(2 * n), yv, L1, L2, o.MtM)
    assert np.allclose(AtKx, ref, atol=1e-12)


@pytest.mark.parametrize("dims", [(64,), (12, 10), (6, 5, 4)])
def test_ipm_matches_fft_model(dims):
    import csipm.workloads as wl
    p = wl.batch_instance(dims, seed=4)
    o = gz.GondzioOps(dims, p.missing)
    res = gz.solve(o, p.z0, p.lam, tol=1e-10)
    assert res.status == "solved"
    b = op.m_perpt_z(p.z0, dims)
    mt = lambda v: op.mtm(v, dims, p.missing)    # noqa: E731
    ref = ipm.solve(b, p.lam, mt, ipm.IPMOptions(tol=1e-10, cg_rtol=1e-13))
    f_ref = ipm.objective(ref.beta, np.abs(ref.beta), mt(ref.beta), b, p.lam) + \
        0.5 * float(p.z0 @ p.z0)
    assert abs(res.objective - f_ref) <= 1e-7 * abs(f_ref)
    n = len(p.z0)
    assert (res.x[:2 * n] > 0).all()
    # planted: a mask-free M^T M (wrong operator) gives another objective
    bad = gz.GondzioOps(dims, None)
    rb = gz.solve(bad, p.z0, p.lam, tol=1e-8)
    fb = gz.obj(np.concatenate([rb.x[:2 * n], p.z0 - o.M(rb.x[:n] - rb.x[n:2 * n])]),
                p.lam, n)
    assert abs(fb - f_ref) > 1e-6 * abs(f_ref)
