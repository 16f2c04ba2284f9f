"""GPU tests of the batch engine (csipm_batch_*).

* A batch of uneven 1D / 2D / 3D problems gives, instance by instance, the
  result of the same problem solved alone (csipm_solve, a batch of one):
  objective within 1e-9 relative, beta within 1e-6 relative l2, the same
  support above 1e-6 max|beta| on entries clearly away from the threshold.  (A batch tiles its
  FFT passes differently, so the CG dot products round differently.)
* Continuous batching: problems admitted while others are mid-solve, and
  retired as they finish, still give the alone results.
* The mixed-precision batch reaches the FP64 oracle objective within 1e-6.
* The CG chunks replayed as captured HIP graphs give bit for bit the result
  of the same launches issued one by one (CSIPM_NO_GRAPH), over the many
  replays of a solve with host synchronisations between them.
"""
import numpy as np
import pytest
import torch

from oracle import ipm
from oracle import operator as op

pytestmark = pytest.mark.gpu


def _problems():
    import csipm.workloads as wl
    return [wl.reference_example((128,), False),
            wl.batch_instance((1500,), seed=21),
            wl.batch_instance((63, 63), seed=22),
            wl.batch_instance((26, 21, 15), seed=23),
            wl.batch_instance((61, 53), seed=24),
            wl.batch_instance((517,), seed=25),
            wl.reference_example((10, 8, 12), False),
            wl.batch_instance((20000,), seed=26),
            wl.batch_instance((12289,), seed=27)]


def _op(p, mixed):
    from csipm.fft_operator import FFTOperator
    return FFTOperator(p.N, len(p.dims), p.dims, p.missing, mixed=mixed)


def _same(r, a, name):
    assert r.status == "solved" and a.status == "solved", name
    assert abs(r.objective - a.objective) <= 1e-9 * abs(a.objective), name
    x = r.beta.cpu().numpy()
    y = a.solution.cpu().numpy()
    assert np.linalg.norm(x - y) <= 1e-6 * np.linalg.norm(y), name
    thr = 1e-6 * np.abs(y).max()
    clear = ((np.abs(y) > 10 * thr) | (np.abs(y) < 0.1 * thr)) & \
        ((np.abs(x) > 10 * thr) | (np.abs(x) < 0.1 * thr))
    assert np.array_equal((np.abs(x) > thr)[clear], (np.abs(y) > thr)[clear]), name


def _alone(p, mixed, tol):
    from csipm.fft_operator import FFTParameters
    from csipm.model import FFTNLPModel, ipm_solve
    params = FFTParameters(len(p.dims), p.dims, p.z0, p.lam, p.missing)
    return ipm_solve(FFTNLPModel(params, mixed=mixed), tol=tol)


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_equals_alone(cuda, mixed):
    from csipm.batch import solve_batch
    print("batch vs alone", "mixed" if mixed else "f64", flush=True)
    tol = 1e-7 if mixed else 1e-8
    probs = _problems()
    ops = [_op(p, mixed) for p in probs]
    res = solve_batch(ops, [p.z0 for p in probs], [p.lam for p in probs], mixed=mixed, tol=tol)
    for p, r in zip(probs, res):
        _same(r, _alone(p, mixed, tol), p.name)


def test_continuous_batching(cuda):
    from csipm.batch import BatchSolver, slots_of
    print("continuous batching", flush=True)
    probs = _problems()
    ops = [_op(p, True) for p in probs]
    # a pool too small for all at once: later problems wait for slots
    cap = sum(sorted(slots_of(o) for o in ops)[-3:])
    bs = BatchSolver(cap, max_instances=4, mixed=True)
    queue = list(range(len(probs)))
    done = {}
    steps = 0
    while queue or bs.live:
        while queue:
            k = queue[0]
            if bs.admit(ops[k], probs[k].z0, probs[k].lam, tag=k) is None:
                break
            queue.pop(0)
        assert bs.live > 0
        bs.step(tol=1e-7)
        steps += 1
        for i in bs.finished():
            r = bs.harvest(i)
            done[r.tag] = r
        assert steps < 2000
    bs.close()
    assert sorted(done) == list(range(len(probs)))
    for k, p in enumerate(probs):
        _same(done[k], _alone(p, True, 1e-7), p.name)


def test_mixed_batch_vs_oracle(cuda):
    from csipm.batch import solve_batch
    print("mixed batch vs oracle", flush=True)
    probs = _problems()
    ops = [_op(p, True) for p in probs]
    res = solve_batch(ops, [p.z0 for p in probs], [p.lam for p in probs], mixed=True, tol=1e-7)
    for p, r in zip(probs, res):
        b = op.m_perpt_z(p.z0, p.dims)
        mt = lambda v: op.mtm(v, p.dims, p.missing)    # noqa: E731
        ref = ipm.solve(b, p.lam, mt, ipm.IPMOptions(tol=1e-10, cg_rtol=1e-13))
        beta = r.beta.cpu().numpy()
        f = ipm.objective(beta, np.abs(beta), mt(beta), b, p.lam)
        f_ref = ipm.objective(ref.beta, np.abs(ref.beta), mt(ref.beta), b, p.lam)
        assert abs(f - f_ref) <= 1e-6 * abs(f_ref), p.name


@pytest.mark.parametrize("mixed", [False, True])
def test_graphs_match_eager(cuda, monkeypatch, mixed):
    from csipm.batch import solve_batch
    print("graphs vs eager", "mixed" if mixed else "f64", flush=True)
    probs = _problems()[:5]
    ops = [_op(p, mixed) for p in probs]
    args = ([p.z0 for p in probs], [p.lam for p in probs])
    tol = 1e-7 if mixed else 1e-8
    res_g = solve_batch(ops, *args, mixed=mixed, tol=tol)
    monkeypatch.setenv("CSIPM_NO_GRAPH", "1")
    res_e = solve_batch(ops, *args, mixed=mixed, tol=tol)
    for p, a, b in zip(probs, res_g, res_e):
        assert a.iterations == b.iterations and a.cg_iterations == b.cg_iterations, p.name
        assert a.objective == b.objective, p.name
        assert torch.equal(a.beta, b.beta), p.name


def test_readmission_after_destroy(cuda):
    """A handle destroyed after its instance retired, and a new one admitted
    into the same id and slots (its address may be the old one's): the
    launch tables are rebuilt for the new admission, so the result is the new
    problem's own (an engine keyed its tables on the handle address once and
    ran a re-admitted instance on the freed handle's buffers)."""
    import gc
    from csipm.batch import BatchSolver, slots_of
    print("readmission", flush=True)
    import csipm.workloads as wl
    pa = wl.batch_instance((237, 237), seed=31)
    pb = wl.batch_instance((237, 237), seed=32)
    op_a = _op(pa, True)
    bs = BatchSolver(slots_of(op_a), max_instances=1, mixed=True)
    for p, make in ((pa, None), (pb, True)):
        op = op_a if make is None else _op(p, True)
        assert bs.admit(op, p.z0, p.lam, tag=0) is not None
        steps = 0
        while bs.live:
            bs.step(tol=1e-7)
            steps += 1
            for i in bs.finished():
                r = bs.harvest(i)
            assert steps < 300
        _same(r, _alone(p, True, 1e-7), "readmitted" if make else "first")
        if make is None:
            del op, op_a
            gc.collect()
            spacer = torch.empty(1 << 22, device=cuda)    # moves later allocations
    bs.close()
    del spacer


def test_serve_stream(cuda):
    """csipm.batch.serve over a short arrival stream (handles made by helper
    threads): every request solved, each result the same as alone."""
    from csipm.batch import Request, serve
    print("serve stream", flush=True)
    probs = _problems()[:6]
    reqs = [Request(0.002 * k, k, p) for k, p in enumerate(probs)]
    out = serve(reqs, lambda p: _op(p, True), mixed=True, pool_budget_bytes=1 << 30,
                tol=1e-7, prep_workers=2)
    assert sorted(r.tag.seq for r in out) == list(range(len(probs)))
    for r in out:
        p = probs[r.tag.seq]
        _same(r, _alone(p, True, 1e-7), p.name)
