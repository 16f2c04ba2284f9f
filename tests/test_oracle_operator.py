"""Pins the CPU operator / KKT / PCG / IPM oracle.

Known answers from the reference's own tests (test/unit_tests_ipm.jl): with no
missing sample M^T M = I (:7-8, :21-22), M_perpt_z and M_perp_beta preserve
norms (:10-14), and the package composition equals the independent fft_wei.jl
composition (:21-30).  Beyond them: a brute-force DFT-matrix build of M on
tiny shapes, exact identities of the KKT algebra against dense solves, CG on
dense SPD systems, the reference's full-solve known answer
(test/ipm_example_1D.jl:74-84) and an algorithm-free LASSO certificate.
"""
import numpy as np
import pytest

from oracle import ipm, kkt, maps
from oracle import operator as op

import csipm.workloads as wl

UNIT = [(100,), (200,), (500,), (16, 16), (12, 18), (16, 24), (32, 32), (8, 8, 8),
        (2, 4, 6), (6, 10, 12), (14, 8, 4), (8, 6, 4), (16, 16, 16)]
ODD = [(9,), (255,), (15, 7), (5, 7, 9), (9, 10, 7)]


@pytest.mark.parametrize("dims", UNIT + ODD)
def test_unit_tests_ipm_identities(dims):
    rng = np.random.default_rng(7)
    n = int(np.prod(dims))
    z = rng.uniform(size=n)
    assert np.allclose(op.mtm(z, dims), z, rtol=0, atol=1e-13 * np.abs(z).max() * 10)
    assert abs(np.linalg.norm(op.m_perpt_z(z, dims)) - np.linalg.norm(z)) < 1e-12 * np.linalg.norm(z)
    assert abs(np.linalg.norm(op.m_perp_beta(z, dims)) - np.linalg.norm(z)) < 1e-12 * np.linalg.norm(z)
    if all(d % 2 == 0 for d in dims):
        np.testing.assert_allclose(op.m_perpt_z(z, dims), op.m_perpt_z_wei(z, dims), atol=1e-13)
        np.testing.assert_allclose(op.m_perp_beta(z, dims), op.m_perp_beta_wei(z, dims), atol=1e-13)


@pytest.mark.parametrize("dims", [(16, 24), (8, 6, 4), (100,)])
def test_masked_matches_wei(dims):
    rng = np.random.default_rng(8)
    n = int(np.prod(dims))
    miss = rng.choice(n, size=n // 7, replace=False)
    x = rng.standard_normal(n)
    np.testing.assert_allclose(op.mtm(x, dims, miss), op.mtm_wei(x, dims, miss), atol=1e-13)


@pytest.mark.parametrize("dims", [(6,), (7,), (4, 6), (5, 3), (3, 4, 5)])
def test_brute_force_dft_matrix(dims):
    """M = D * Re(F^{-1}) * B built from explicit DFT matrices and the map
    table; compared with the FFT-based oracle column by column."""
    n = int(np.prod(dims))
    rng = np.random.default_rng(9)
    miss = rng.choice(n, size=max(1, n // 5), replace=False)
    # explicit unitary inverse DFT on the Fortran-flattened multi-index
    coords = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), -1)
    coords = coords.reshape(-1, len(dims), order="F")
    phase = np.zeros((n, n))
    for ax, d in enumerate(dims):
        phase += np.outer(coords[:, ax], coords[:, ax]) / d
    Finv = np.exp(2j * np.pi * phase) / np.sqrt(n)
    B = np.stack([maps.beta_to_dft(e, dims, rdft=False) for e in np.eye(n)], axis=1)
    Mdense = (Finv @ B).real
    Mdense[miss] = 0.0
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        np.testing.assert_allclose(op.m_perp_beta(e, dims, miss), Mdense[:, i], atol=1e-14)
    Q = op.dense_mtm(dims, miss)
    np.testing.assert_allclose(Q, Mdense.T @ Mdense, atol=1e-14)
    assert np.linalg.eigvalsh(0.5 * (Q + Q.T)).max() <= 1 + 1e-12


def test_kkt_algebra_dense():
    dims = (4, 6)
    n = 24
    rng = np.random.default_rng(10)
    miss = rng.choice(n, size=5, replace=False)
    Q = op.dense_mtm(dims, miss)
    s1 = rng.uniform(0.1, 3, n)
    s2 = rng.uniform(0.1, 3, n)
    L1, L2, P11, P12, P22 = kkt.precond_build(s1, s2)
    K = np.block([[Q + np.diag(L1), np.diag(L2)], [np.diag(L2), np.diag(L1)]])
    x = rng.standard_normal(2 * n)
    y0 = rng.standard_normal(2 * n)
    got = kkt.kkt_mul(y0, x, L1, L2, lambda v: op.mtm(v, dims, miss), 0.7, -0.3)
    np.testing.assert_allclose(got, 0.7 * K @ x - 0.3 * y0, atol=1e-13)
    # preconditioner = exact inverse with M^T M replaced by I
    Kp = np.block([[np.eye(n) + np.diag(L1), np.diag(L2)], [np.diag(L2), np.diag(L1)]])
    P = np.block([[np.diag(P11), np.diag(P12)], [np.diag(P12), np.diag(P22)]])
    np.testing.assert_allclose(P @ Kp, np.eye(2 * n), atol=1e-12)
    np.testing.assert_allclose(kkt.precond_mul(y0, x, P11, P12, P22, 2.0, 0.5),
                               2.0 * P @ x + 0.5 * y0, atol=1e-13)
    # condensation + unpack solve the 6N KKT system [[H, J^T], [J, 0]]
    H = np.zeros((4 * n, 4 * n))
    H[:n, :n] = Q
    H[2 * n:3 * n, 2 * n:3 * n] = np.diag(s1)
    H[3 * n:, 3 * n:] = np.diag(s2)
    I = np.eye(n)
    Z = np.zeros((n, n))
    J = np.block([[-I, -I, -I, Z], [I, -I, Z, -I]])
    A = np.block([[H, J.T], [J, np.zeros((2 * n, 2 * n))]])
    w = rng.standard_normal(6 * n)
    sol_dense = np.linalg.solve(A, w)
    b = kkt.condense_rhs(w, s1, s2)
    xr = np.linalg.solve(K, b)
    sol = kkt.unpack_solution(w, xr, s1, s2)
    np.testing.assert_allclose(sol, sol_dense, atol=1e-10)
    # full-KKT product and J^T
    mt = lambda v: op.mtm(v, dims, miss)  # noqa: E731
    v6 = rng.standard_normal(6 * n)
    A0 = A.copy()
    A0[2 * n:4 * n, 2 * n:4 * n] = 0.0           # _kktmul! adds the diagonals
    np.testing.assert_allclose(kkt.full_kkt_mul(v6, mt, n), A0 @ v6, atol=1e-13)
    y2 = rng.standard_normal(2 * n)
    np.testing.assert_allclose(kkt.jtprod(y2, n), J.T @ y2, atol=1e-14)


def test_pcg_dense():
    rng = np.random.default_rng(11)
    n = 40
    G = rng.standard_normal((n, n))
    A = G @ G.T + n * np.eye(n)
    Dinv = 1.0 / np.diag(A)
    b = rng.standard_normal(n)
    x, it, res, st = kkt.pcg(lambda v: A @ v, b, lambda r: Dinv * r, atol=1e-12, rtol=0.0)
    assert st == "solved" and it <= n
    np.testing.assert_allclose(x, np.linalg.solve(A, b), atol=1e-10)
    # zero curvature
    x, it, res, st = kkt.pcg(lambda v: -v, b, None, atol=1e-12)
    assert st == "zero-curvature" and it == 0


@pytest.mark.parametrize("dims", [(100,), (200,), (16, 16), (12, 18), (8, 8, 8), (2, 4, 6),
                                  (6, 10, 12)])
def test_ipm_known_answer(dims):
    """ipm_example_*D(check=true): lambda = 0, no mask, no noise -> beta equals
    DFT_to_beta(fft(x)/sqrt(N)) within 1e-6 (test/ipm_example_1D.jl:74-84)."""
    import csipm.workloads as wl
    p = wl.reference_example(dims, check=True)
    b = op.m_perpt_z(p.z0, dims)
    r = ipm.solve(b, p.lam, lambda v: op.mtm(v, dims, p.missing))
    assert r.status == "solved"
    n = int(np.prod(dims))
    vh = np.fft.rfftn(p.x_true.reshape(dims[::-1])).reshape(-1) / np.sqrt(n)
    beta_true = maps.dft_to_beta(vh, dims)
    assert np.linalg.norm(beta_true - r.beta) <= 1e-6


@pytest.mark.parametrize("dims", [(100,), (16, 16), (8, 8, 8)])
def test_ipm_lasso_certificate(dims):
    import csipm.workloads as wl
    p = wl.reference_example(dims, check=False)
    b = op.m_perpt_z(p.z0, dims)
    mt = lambda v: op.mtm(v, dims, p.missing)   # noqa: E731
    r = ipm.solve(b, p.lam, mt, ipm.IPMOptions(tol=1e-10, cg_rtol=1e-13))
    assert r.status == "solved"
    fp, gap, primal = ipm.lasso_certificate(r.beta, b, p.lam, mt,
                                            lambda v: op.m_perp_beta(v, dims, p.missing), p.z0)
    assert gap >= -1e-9 * abs(primal)
    assert gap <= 1e-8 * abs(primal)
    # objective of the reference model at c = |beta| equals P - 1/2||z0||^2
    f = ipm.objective(r.beta, np.abs(r.beta), mt(r.beta), b, p.lam)
    assert abs(f + 0.5 * p.z0 @ p.z0 - primal) <= 1e-9 * abs(primal)


def test_ipm_planted_operator_mistake():
    """A dropped term (the sampling mask left out of M^T M, i.e. Q = I) yields a
    different minimiser; the certificate computed with the right operator
    flags it."""
    import csipm.workloads as wl
    dims = (64,)
    p = wl.reference_example(dims, check=False)
    b = op.m_perpt_z(p.z0, dims)
    right = lambda v: op.mtm(v, dims, p.missing)      # noqa: E731
    wrong = lambda v: v.copy()                        # noqa: E731
    r = ipm.solve(b, p.lam, wrong, ipm.IPMOptions(tol=1e-10, cg_rtol=1e-13))
    assert r.status == "solved"
    fp, gap, primal = ipm.lasso_certificate(r.beta, b, p.lam, right,
                                            lambda v: op.m_perp_beta(v, dims, p.missing), p.z0)
    assert gap > 1e-4 * abs(primal)
    r = ipm.solve(b, p.lam, right, ipm.IPMOptions(tol=1e-10, cg_rtol=1e-13))
    fp, gap, primal = ipm.lasso_certificate(r.beta, b, p.lam, right,
                                            lambda v: op.m_perp_beta(v, dims, p.missing), p.z0)
    assert gap < 1e-8 * abs(primal)


@pytest.mark.parametrize("dims", [(100,), (16, 12), (9, 10, 7), (15, 17), (8, 6, 4)])
def test_complex_path_equals_real_path(dims):
    """rdft=False (fft_utils.jl:80-81,109-110: complex FFTs over the full
    Hermitian spectrum) against the real-to-complex statement and the
    independent full-spectrum block maps of test/fft_wei.jl."""
    n = int(np.prod(dims))
    rng = np.random.default_rng(n + 1)
    x = rng.standard_normal(n)
    miss = np.sort(rng.choice(n, size=n // 6, replace=False))
    c = op.mtm(x, dims, miss, rdft=False)
    assert np.abs(c - op.mtm(x, dims, miss)).max() <= 1e-13 * np.abs(c).max()
    if all(d % 2 == 0 for d in dims):
        assert np.abs(c - op.mtm_wei(x, dims, miss)).max() <= 1e-13 * np.abs(c).max()
    y = op.m_perp_beta(x, dims, miss, rdft=False)
    assert np.abs(y - op.m_perp_beta(x, dims, miss)).max() <= 1e-13 * np.abs(y).max()
    b = op.m_perpt_z(x, dims, rdft=False)
    assert np.abs(b - op.m_perpt_z(x, dims)).max() <= 1e-13 * np.abs(b).max()


@pytest.mark.parametrize("dims", [(100,), (200,), (16, 16), (12, 18), (8, 8, 8), (2, 4, 6)])
def test_admm_known_answer(dims):
    """test/admm_example_*D.jl with check=true: lambda = 0, no missing
    samples, rho = 1, tol 1e-8 -> beta = DFT_to_beta(fft(x)/sqrt(N)) within 1e-6."""
    from oracle import admm
    p = wl.reference_example(dims, check=True)
    b = op.m_perpt_z(p.z0, dims)
    res = admm.solve(b, 0.0, lambda v: op.mtm(v, dims), rho=1.0, maxt=1000, tol=1e-8)
    assert res.converged
    n = p.N
    vh = np.fft.rfftn(p.x_true.reshape(dims[::-1])).reshape(-1) / np.sqrt(n)
    assert np.linalg.norm(maps.dft_to_beta(vh, dims) - res.beta) <= 1e-6


def test_admm_solves_the_lasso():
    """On a masked problem with lambda > 0 the ADMM fixed point is the LASSO
    minimiser: its duality gap matches the IPM's certified optimum."""
    from oracle import admm
    p = wl.batch_instance((96,), seed=3)
    dims = p.dims
    b = op.m_perpt_z(p.z0, dims)
    mt = lambda v: op.mtm(v, dims, p.missing)          # noqa: E731
    res = admm.solve(b, p.lam, mt, rho=0.1, maxt=3000, tol=1e-9)
    assert res.converged
    fp, gap, primal = ipm.lasso_certificate(res.beta, b, p.lam, mt,
                                            lambda v: op.m_perp_beta(v, dims, p.missing), p.z0)
    assert gap <= 1e-8 * abs(primal) and fp <= 1e-9
    # a planted mistake (threshold at lambda instead of lambda / rho, rho = 2) is caught
    z = res.beta
    bad = admm.soft_threshold(z + 0.1, p.lam) - admm.soft_threshold(z + 0.1, p.lam / 2.0)
    assert np.abs(bad).max() > 0
