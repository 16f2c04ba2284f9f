"""GPU parity of the Gondzio formulation (src/gondzio_model.jl, gondzio_kkt.jl).

* obj / grad / cons / jtprod, the condensed K, the full-KKT product, the
  preconditioner, condense / unpack against the CPU oracle (1e-12 relative).
* The condensed-K PCG (run through the congruent FFT system): iteration
  count within 1 of the oracle PCG on the Gondzio K itself, solution within
  1e-9; solve_kkt of GondzioKKTSystem against the oracle pipeline.
* gondzio_solve: objective at (beta+, beta-, z0 - M beta) within 1e-6 of
  the oracle's Gondzio interior-point solve (the reported interior-point
  objective within 1e-4: it also holds the final barrier gap), p, q >= 0,
  y = r, c(x) = 0 to rounding; mixed precision likewise.
"""
import numpy as np
import pytest
import torch

from oracle import gondzio as gz

pytestmark = pytest.mark.gpu

SHAPES = [(100,), (8192,), (31, 17), (64, 48), (12, 10, 8)]


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _model(p, mixed=False):
    from csipm.fft_operator import FFTParameters
    from csipm.gondzio import GondzioNLPModel
    params = FFTParameters(len(p.dims), p.dims, p.z0, p.lam, p.missing)
    return GondzioNLPModel(params, mixed=mixed)


@pytest.mark.parametrize("dims", SHAPES)
def test_model_and_kkt_pieces(cuda, dims):
    import csipm.workloads as wl
    from csipm.gondzio import GondzioKKTSystem
    print("gondzio pieces", dims, flush=True)
    p = wl.batch_instance(dims, seed=7)
    n = int(np.prod(dims))
    o = gz.GondzioOps(dims, p.missing)
    nlp = _model(p)
    kk = GondzioKKTSystem(nlp)
    rng = np.random.default_rng(5)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)    # noqa: E731
    x3 = np.abs(rng.standard_normal(3 * n))
    assert abs(nlp.obj(t(x3)) - gz.obj(x3, p.lam, n)) <= 1e-12 * gz.obj(x3, p.lam, n)
    assert _rel(nlp.grad(t(x3)).cpu(), gz.grad(x3, p.lam, n)) <= 1e-15
    assert _rel(nlp.cons(t(x3)).cpu(), gz.cons(o, x3, p.z0)) <= 1e-12
    v = rng.standard_normal(n)
    assert _rel(kk.jtprod(t(v)).cpu(), gz.jtprod(o, v)) <= 1e-12
    x4 = rng.standard_normal(4 * n)
    y4 = rng.standard_normal(4 * n)
    assert _rel(kk.mul(t(x4), 1.3, 0.5, t(y4.copy())).cpu(),
                gz.kkt_mul_full(o, x4, y4, 1.3, 0.5)) <= 1e-12
    Wp, Wq = rng.random(n) + 0.05, rng.random(n) + 0.05
    kk.pr_diag[:n] = t(Wp)
    kk.pr_diag[n:2 * n] = t(Wq)
    kk.build_kkt()
    P = gz.precond_build(Wp, Wq)
    for got, want in zip((kk.P.P11, kk.P.P12, kk.P.P22), P):
        assert _rel(got.cpu(), want) <= 1e-12
    x2 = rng.standard_normal(2 * n)
    y2 = rng.standard_normal(2 * n)
    assert _rel(kk.K.mul(t(x2), 0.7, -0.3, t(y2.copy())).cpu(),
                gz.kkt_mul(o, x2, y2, Wp, Wq, 0.7, -0.3)) <= 1e-12
    assert _rel(kk.P.mul(t(x2)).cpu(), gz.precond_apply(x2, np.zeros(2 * n), *P, 1.0, 0.0)) \
        <= 1e-12
    w = rng.standard_normal(4 * n)
    rhs = gz.condense_rhs(o, w)
    dx_ref, it_ref, _, st = gz.pcg(o, rhs, Wp, Wq, P, atol=1e-12)
    got = kk.solve_kkt(t(w.copy())).cpu().numpy()
    want = gz.unpack(o, w, dx_ref)
    assert abs(kk.krylov_iterations[-1] - it_ref) <= 1
    assert _rel(got, want) <= 1e-9


@pytest.mark.parametrize("dims", [(200,), (5000,), (63, 63), (26, 21, 15)])
@pytest.mark.parametrize("mixed", [False, True])
def test_gondzio_solve(cuda, dims, mixed):
    import csipm.workloads as wl
    from csipm.gondzio import gondzio_solve
    print("gondzio solve", dims, "mixed" if mixed else "f64", flush=True)
    p = wl.batch_instance(dims, seed=8)
    n = int(np.prod(dims))
    nlp = _model(p, mixed=mixed)
    st = gondzio_solve(nlp, tol=1e-8)
    assert st.status == "solved"
    o = gz.GondzioOps(dims, p.missing)
    ref = gz.solve(o, p.z0, p.lam, tol=1e-10)
    x = st.solution.cpu().numpy()

    def clean(xx):
        # the objective at p = beta+, q = beta-: an interior iterate also
        # carries lambda sum(p + q - |beta|), of the order N mu / lambda
        b = xx[:n] - xx[n:2 * n]
        return gz.obj(np.concatenate([np.maximum(b, 0), np.maximum(-b, 0),
                                      p.z0 - o.M(b)]), p.lam, n)
    assert abs(clean(x) - clean(ref.x)) <= 1e-6 * abs(clean(ref.x))
    assert abs(st.objective - ref.objective) <= 1e-4 * abs(ref.objective)
    assert (x[:2 * n] >= 0).all()
    assert np.array_equal(st.multipliers.cpu().numpy(), x[2 * n:])
    c = gz.cons(o, x, p.z0)
    assert np.abs(c).max() <= 1e-12 * max(1.0, np.abs(p.z0).max())
    assert abs(gz.obj(x, p.lam, n) - st.objective) <= 1e-12 * abs(st.objective)
