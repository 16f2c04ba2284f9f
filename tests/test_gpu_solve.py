"""GPU parity of the NLP model, the full KKT system and the full solve.

* FFTNLPModel obj / grad / cons, FFTKKTSystem mul / jtprod / solve_kkt against
  the CPU oracle restatements of src/fft_model.jl and src/fft_kkt.jl (1e-12).
* The reference's full-solve known answer (test/ipm_example_*D.jl, check=true:
  lambda = 0, no mask) within 1e-6.
* Masked problems, FP64 and mixed precision at tol 1e-8: objective within
  1e-6 relative of the FP64 CPU oracle solved to 1e-10 (and certified by the
  LASSO duality gap); the same support above 1e-6 * max|beta| as the oracle
  solved with the same settings (tol 1e-8), every flip listed (none exempt).
"""
import numpy as np
import pytest
import torch

import _checks as chk
from oracle import ipm, kkt, maps
from oracle import operator as op

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _model(p, dtype=torch.float64, mixed=False):
    from csipm.fft_operator import FFTParameters
    from csipm.model import FFTNLPModel
    params = FFTParameters(len(p.dims), p.dims, p.z0, p.lam, p.missing)
    return FFTNLPModel(params, rdft=True, dtype=dtype, mixed=mixed)


@pytest.mark.parametrize("dims", [(100,), (255, 255), (16, 12, 10)])
def test_model_and_kkt_system(cuda, dims):
    import csipm.workloads as wl
    from csipm.model import FFTKKTSystem
    print("model", dims, flush=True)
    p = wl.batch_instance(dims, seed=5) if len(dims) < 3 else wl.reference_example(dims, False)
    nlp = _model(p)
    n = nlp.nbeta
    mt = lambda v: op.mtm(v, p.dims, p.missing)    # noqa: E731
    b = op.m_perpt_z(p.z0, p.dims)
    assert _rel(nlp.M_perpt_z0.cpu().numpy(), b) <= 1e-12
    rng = np.random.default_rng(1)
    x = rng.standard_normal(2 * n)
    xt = torch.from_numpy(x).to(cuda)
    f_ref = 0.5 * float(op.m_perp_beta(x[:n], p.dims, p.missing) @ op.m_perp_beta(x[:n], p.dims, p.missing)) \
        - float(x[:n] @ b) + p.lam * float(x[n:].sum())
    assert abs(nlp.obj(xt) - f_ref) <= 1e-12 * max(1.0, abs(f_ref)) * 10
    g = nlp.grad(xt).cpu().numpy()
    assert _rel(g[:n], mt(x[:n]) - b) <= 1e-12
    assert np.all(g[n:] == p.lam)
    c = nlp.cons(xt).cpu().numpy()
    assert np.array_equal(c, np.concatenate([-x[:n] - x[n:], x[:n] - x[n:]]))
    # full KKT system pieces
    K = FFTKKTSystem(nlp)
    v6 = torch.from_numpy(rng.standard_normal(6 * n)).to(cuda)
    y6 = torch.from_numpy(rng.standard_normal(6 * n)).to(cuda)
    want = kkt.full_kkt_mul(v6.cpu().numpy(), mt, n, 0.5, 2.0, y6.cpu().numpy())
    assert _rel(K.mul(y6, v6, 0.5, 2.0).cpu().numpy(), want) <= 1e-12
    yy = torch.from_numpy(rng.standard_normal(2 * n)).to(cuda)
    out = torch.empty(4 * n, dtype=torch.float64, device=cuda)
    assert np.array_equal(K.jtprod(out, yy).cpu().numpy(), kkt.jtprod(yy.cpu().numpy(), n))
    # build_kkt + solve_kkt against a CPU solve of the same 6N system
    s1 = rng.uniform(0.1, 4.0, n)
    s2 = rng.uniform(0.1, 4.0, n)
    K.pr_diag[2 * n:3 * n] = torch.from_numpy(s1).to(cuda)
    K.pr_diag[3 * n:] = torch.from_numpy(s2).to(cuda)
    K.build_kkt()
    w = rng.standard_normal(6 * n)
    wt = torch.from_numpy(w).to(cuda)
    K.solve_kkt(wt)
    L1, L2, P11, P12, P22 = kkt.precond_build(s1, s2)
    rhs = kkt.condense_rhs(w, s1, s2)
    xr, *_ = kkt.pcg(lambda v: kkt.kkt_mul(np.zeros(2 * n), v, L1, L2, mt), rhs,
                     lambda r: kkt.precond_mul(np.zeros(2 * n), r, P11, P12, P22), atol=1e-12)
    want = kkt.unpack_solution(w, xr, s1, s2)
    assert K.last_stats.status == "solved"
    assert _rel(wt.cpu().numpy(), want) <= 1e-9


@pytest.mark.parametrize("dims", [(100,), (200,), (16, 16), (12, 18), (8, 8, 8), (2, 4, 6),
                                  (6, 10, 12)])
def test_known_answer(cuda, dims):
    import csipm.workloads as wl
    from csipm.model import ipm_solve
    print("known answer", dims, flush=True)
    p = wl.reference_example(dims, check=True)
    st = ipm_solve(_model(p))
    assert st.status == "solved"
    n = p.N
    vh = np.fft.rfftn(p.x_true.reshape(dims[::-1])).reshape(-1) / np.sqrt(n)
    beta_true = maps.dft_to_beta(vh, dims)
    assert np.linalg.norm(beta_true - st.solution.cpu().numpy()) <= 1e-6


def _support(beta):
    return np.abs(beta) > 1e-6 * np.abs(beta).max()


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("case", ["ex1d", "ex2d", "ex3d", "s1d", "b1d", "b2d", "b3d", "p1d",
                                  "p2d", "bs2d", "r3d", "big1d", "bigp1d"])
def test_solve_vs_oracle(cuda, case, mixed):
    import csipm.workloads as wl
    from csipm.model import ipm_solve
    print("solve", case, "mixed" if mixed else "f64", flush=True)
    p = {"ex1d": lambda: wl.reference_example((128,), False),
         "ex2d": lambda: wl.reference_example((24, 20), False),
         "ex3d": lambda: wl.reference_example((10, 8, 12), False),
         "s1d": lambda: wl.signal_1d_1024(),
         "b1d": lambda: wl.batch_instance((1500,), seed=11),
         "b2d": lambda: wl.batch_instance((63, 63), seed=12),
         "b3d": lambda: wl.batch_instance((26, 21, 15), seed=13),
         "p1d": lambda: wl.batch_instance((2003,), seed=14),
         "p2d": lambda: wl.batch_instance((61, 53), seed=15),
         "bs2d": lambda: wl.batch_instance((83, 111), seed=18),
         "r3d": lambda: wl.batch_instance((37, 23, 19), seed=19),
         "big1d": lambda: wl.batch_instance((20000,), seed=16),
         "bigp1d": lambda: wl.batch_instance((12289,), seed=17)}[case]()
    b = op.m_perpt_z(p.z0, p.dims)
    mt = lambda v: op.mtm(v, p.dims, p.missing)    # noqa: E731
    ref = ipm.solve(b, p.lam, mt, ipm.IPMOptions(tol=1e-10, cg_rtol=1e-13))
    assert ref.status == "solved"
    fp, gap, primal = ipm.lasso_certificate(ref.beta, b, p.lam, mt,
                                            lambda v: op.m_perp_beta(v, p.dims, p.missing), p.z0)
    assert gap <= 1e-8 * abs(primal)
    st = ipm_solve(_model(p, mixed=mixed), tol=1e-8)
    assert st.status == "solved"
    beta = st.solution.cpu().numpy()
    f_gpu = ipm.objective(beta, np.abs(beta), mt(beta), b, p.lam)
    f_ref = ipm.objective(ref.beta, np.abs(ref.beta), mt(ref.beta), b, p.lam)
    assert abs(f_gpu - f_ref) <= 1e-6 * abs(f_ref)
    # support: against the oracle with the same settings (an interior-point
    # iterate at tol 1e-8 can still hold an off-support entry just above
    # 1e-6 max|beta|: the oracle at 1e-8 and at 1e-10 differ so on big1d)
    ref8 = ipm.solve(b, p.lam, mt, ipm.IPMOptions(tol=1e-8))
    chk.assert_same_support(beta, ref8.beta, f"{case} {'mixed' if mixed else 'f64'}")


def test_default_model_is_complex_path(cuda):
    """FFTNLPModel defaults to rdft=False like the reference (fft_model.jl:14);
    the solve reaches the oracle objective within 1e-6."""
    import csipm.workloads as wl
    from csipm.fft_operator import FFTParameters
    from csipm.model import FFTNLPModel, ipm_solve
    p = wl.batch_instance((63, 63), seed=31)
    nlp = FFTNLPModel(FFTParameters(2, p.dims, p.z0, p.lam, p.missing))
    assert nlp.op_fft.rdft is False
    st = ipm_solve(nlp, tol=1e-8)
    b = op.m_perpt_z(p.z0, p.dims)
    mt = lambda v: op.mtm(v, p.dims, p.missing)    # noqa: E731
    ref = ipm.solve(b, p.lam, mt, ipm.IPMOptions(tol=1e-10, cg_rtol=1e-13))
    beta = st.solution.cpu().numpy()
    f = ipm.objective(beta, np.abs(beta), mt(beta), b, p.lam)
    f_ref = ipm.objective(ref.beta, np.abs(ref.beta), mt(ref.beta), b, p.lam)
    assert st.status == "solved" and abs(f - f_ref) <= 1e-6 * abs(f_ref)


@pytest.mark.parametrize("dims", [(100,), (200,), (500,), (16, 16), (12, 18), (16, 24),
                                  (32, 32), (8, 8, 8), (2, 4, 6), (6, 10, 12), (14, 8, 4),
                                  (8, 6, 4), (16, 16, 16)])
def test_admm_known_answer(cuda, dims):
    """test/unit_tests_admm.jl: every shape, check=true (lambda = 0, no mask,
    rho = 1, tol 1e-8) -> beta = DFT_to_beta(fft(x)/sqrt(N)) within 1e-6."""
    import csipm.workloads as wl
    from csipm.admm import CompressedSensingADMM
    from csipm.fft_operator import FFTOperator, FFTParameters
    print("admm known answer", dims, flush=True)
    p = wl.reference_example(dims, check=True)
    params = FFTParameters(len(dims), dims, p.z0, 0.0, p.missing)
    st = CompressedSensingADMM(params, FFTOperator(p.N, len(dims), dims, p.missing), rho=1.0,
                               maxt=1000, tol=1e-8, stats=True)
    assert st.status == "converged"
    vh = np.fft.rfftn(p.x_true.reshape(dims[::-1])).reshape(-1) / np.sqrt(p.N)
    assert np.linalg.norm(maps.dft_to_beta(vh, dims) - st.solution.cpu().numpy()) <= 1e-6


@pytest.mark.parametrize("case", ["1d", "2d", "3d"])
def test_admm_vs_oracle(cuda, case):
    """Masked problems with lambda > 0: the device ADMM follows the oracle's
    iteration (same count within 1) to the same point (1e-8 relative)."""
    import csipm.workloads as wl
    from csipm.admm import CompressedSensingADMM
    from oracle import admm
    from csipm.fft_operator import FFTOperator, FFTParameters
    print("admm vs oracle", case, flush=True)
    p = {"1d": lambda: wl.batch_instance((300,), seed=41),
         "2d": lambda: wl.batch_instance((24, 20), seed=42),
         "3d": lambda: wl.batch_instance((10, 8, 12), seed=43)}[case]()
    b = op.m_perpt_z(p.z0, p.dims)
    mt = lambda v: op.mtm(v, p.dims, p.missing)    # noqa: E731
    ref = admm.solve(b, p.lam, mt, rho=0.1, maxt=3000, tol=1e-8)
    params = FFTParameters(len(p.dims), p.dims, p.z0, p.lam, p.missing)
    st = CompressedSensingADMM(params, FFTOperator(p.N, len(p.dims), p.dims, p.missing),
                               rho=0.1, maxt=3000, tol=1e-8, stats=True)
    assert ref.converged and st.status == "converged"
    assert abs(st.iterations - ref.iterations) <= 1
    z = st.solution.cpu().numpy()
    assert np.linalg.norm(z - ref.beta) <= 1e-8 * np.linalg.norm(ref.beta) + 1e-10
