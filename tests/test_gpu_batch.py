"""GPU tests of the batch engine (csipm_batch_*).

* A batch of uneven 1D / 2D / 3D problems gives, instance by instance, the
  result of the same problem solved alone (csipm_solve, a batch of one):
  objective within 1e-9 relative, beta within 1e-6 relative l2, the same
  support above 1e-6 max|beta| (every flip listed).  (A batch tiles its
  FFT passes differently, so the CG dot products round differently.)
* Admission rules: a handle serves one live instance (admitting it twice
  raises); serve() returns a request that can never fit its budget as
  "rejected" at once and solves the others.  The batched admission /
  harvest calls give per problem what the single ones give.
* Failure statuses forced from the host: an IPM iteration cap, a CG
  iteration cap of one, a non-finite sample in z0.
* Continuous batching: problems admitted while others are mid-solve, and
  retired as they finish, still give the alone results; so do serve()'s
  two-lane form (two engines stepped in parallel threads).
* The mixed-precision batch reaches the FP64 oracle objective within 1e-6.
* The CG chunks replayed as captured HIP graphs give bit for bit the result
  of the same launches issued one by one (CSIPM_NO_GRAPH), over the many
  replays of a solve with host synchronisations between them.
"""
import numpy as np
import pytest
import torch

import _checks as chk
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
    chk.assert_same_support(x, y, name)


def _alone(p, mixed, tol):
    from csipm.fft_operator import FFTParameters
    from csipm.model import FFTNLPModel, ipm_solve
    params = FFTParameters(len(p.dims), p.dims, p.z0, p.lam, p.missing)
    return ipm_solve(FFTNLPModel(params, mixed=mixed), tol=tol)


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_equals_alone(cuda, mixed):
    from csipm.batch import solve_batch
    print("batch vs alone", "mixed" if mixed else "f64", flush=True)
    tol = 1e-8
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


def test_admit_harvest_many(cuda):
    """The batched admission / harvest calls behave as the single ones: per
    problem, an id, None when it does not fit (a later, smaller one is still
    admitted) or the error (a handle given twice); results equal the alone
    solves."""
    from csipm import _lib
    from csipm.batch import BatchSolver, slots_of
    print("admit / harvest many", flush=True)
    probs = _problems()
    ops = [_op(p, True) for p in probs]
    sl = [slots_of(o) for o in ops]
    big = int(np.argmax(sl))
    small = [k for k in range(len(probs)) if k != big]
    # room for every problem but the largest; admitted largest-first except
    # the smallest, the largest then finds only the smallest's room left and
    # does not fit, the smallest after it does; one handle is given twice
    small.sort(key=lambda k: -sl[k])
    bs = BatchSolver(sum(sl[k] for k in small), max_instances=len(probs), mixed=True)
    order = small[:-1] + [big, small[-1], small[0]]
    items = []
    for k in order:
        z, beta = BatchSolver.buffers(ops[k], probs[k].z0)
        items.append((ops[k], z, beta, probs[k].lam, k))
    got = bs.admit_many(items)
    assert all(isinstance(g, int) for g in got[:len(small) - 1]), got
    assert got[-3] is None, got
    assert isinstance(got[-2], int), got
    assert isinstance(got[-1], _lib.CsipmError) and got[-1].code == _lib.ERR_ARG, got
    assert bs.live == len(small)
    done = {}
    while bs.live:
        bs.step(tol=1e-7)
        for r in bs.harvest_many(bs.finished()):
            done[r.tag] = r
    with pytest.raises(_lib.CsipmError):
        bs.harvest_many([0])             # nothing admitted under that id any more
    bs.close()
    assert sorted(done) == sorted(small)
    for k in small:
        _same(done[k], _alone(probs[k], True, 1e-7), probs[k].name)


def test_mixed_batch_vs_oracle(cuda):
    from csipm.batch import solve_batch
    print("mixed batch vs oracle", flush=True)
    probs = _problems()
    ops = [_op(p, True) for p in probs]
    res = solve_batch(ops, [p.z0 for p in probs], [p.lam for p in probs], mixed=True, tol=1e-8)
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
    out = serve(reqs, lambda p: _op(p, True), mixed=True, budget_bytes=1 << 30,
                tol=1e-7, prep_workers=2)
    assert sorted(r.tag.seq for r in out) == list(range(len(probs)))
    for r in out:
        p = probs[r.tag.seq]
        _same(r, _alone(p, True, 1e-7), p.name)


def test_serve_lanes(cuda):
    """serve() over two lanes (a 3D engine and a 1D/2D engine whose
    interior-point iterations run in parallel threads; pools too small for
    all at once): every request solved, each result the same as alone."""
    from csipm.batch import BatchSolver, Request, serve, slots_of
    print("serve lanes", flush=True)
    probs = _problems()
    sl = [slots_of(_op(p, True)) for p in probs]
    three = [k for k, p in enumerate(probs) if len(p.dims) == 3]
    rest = [k for k in range(len(probs)) if k not in three]
    lane3 = BatchSolver(max(sl[k] for k in three), max_instances=2, mixed=True)
    lane_s = BatchSolver(sum(sorted(sl[k] for k in rest)[-2:]), max_instances=3, mixed=True)
    reqs = [Request(0.0, k, p) for k, p in enumerate(probs)]
    st = {}
    out = serve(reqs, lambda p: _op(p, True), mixed=True, budget_bytes=1 << 30, tol=1e-7,
                prep_workers=2, stats=st,
                lanes=[(lane3, lambda p: len(p.dims) == 3), (lane_s, None)])
    lane3.close()
    lane_s.close()
    assert sorted(r.tag.seq for r in out) == list(range(len(probs)))
    assert st["loop_steps"] > 0
    for r in out:
        p = probs[r.tag.seq]
        _same(r, _alone(p, True, 1e-7), p.name)


def test_handle_admitted_once(cuda):
    """One handle per live instance: its passes run in the handle's scratch
    spectrum, so a second admission of a live handle is refused (and so is
    destroying it); after harvest it can be admitted again."""
    import csipm.workloads as wl
    from csipm import _lib
    from csipm.batch import BatchSolver, slots_of
    print("handle admitted once", flush=True)
    p = wl.batch_instance((63, 63), seed=51)
    o = _op(p, True)
    bs = BatchSolver(4 * slots_of(o), max_instances=4, mixed=True)
    assert bs.admit(o, p.z0, p.lam, tag=0) is not None
    with pytest.raises(_lib.CsipmError) as e:
        bs.admit(o, p.z0, p.lam, tag=1)
    assert e.value.code == _lib.ERR_ARG
    assert _lib.lib().csipm_op_destroy(o._h) == _lib.ERR_ARG
    while bs.step(tol=1e-7) > 0:
        pass
    r = bs.harvest(bs.finished()[0])
    assert r.status == "solved"
    assert bs.admit(o, p.z0, p.lam, tag=2) is not None
    bs.close()          # drops the live instance: the handle is free again
    o.close()


def test_serve_rejects_oversize(cuda):
    """A request whose pool slots exceed the serving budget comes back at
    once as "rejected" (a loop that kept it waiting never ended); the other
    requests are still solved, each as alone."""
    import csipm.workloads as wl
    from csipm.batch import Request, footprint, pool_bytes, serve
    print("serve rejects an oversize request", flush=True)
    small = [wl.batch_instance((63, 63), seed=61), wl.batch_instance((1500,), seed=62),
             wl.batch_instance((61, 53), seed=63)]
    big = wl.batch_instance((40, 40, 40), seed=64)
    slots_big, _ = footprint(big.dims, len(big.missing), True)
    slots_small = max(footprint(p.dims, len(p.missing), True)[0] for p in small)
    assert slots_big > 3 * slots_small
    # a pool that holds two small requests but never the big one, beside a
    # handle share of 256 MB (the pool gets pool_fraction of the budget)
    pool = pool_bytes(2 * slots_small, True)
    assert pool < pool_bytes(slots_big, True)
    budget = pool + (256 << 20)
    probs = small[:2] + [big] + small[2:]
    reqs = [Request(0.001 * k, k, p) for k, p in enumerate(probs)]
    t0 = __import__("time").perf_counter()
    out = serve(reqs, lambda p: _op(p, True), mixed=True, budget_bytes=budget,
                pool_fraction=pool / budget, tol=1e-7, prep_workers=2)
    assert __import__("time").perf_counter() - t0 < 60
    got = {r.tag.seq: r for r in out}
    assert sorted(got) == list(range(len(probs)))
    assert got[2].status == "rejected" and "pool slots" in got[2].error
    for k in (0, 1, 3):
        _same(got[k], _alone(probs[k], True, 1e-7), probs[k].name)


def test_failure_statuses(cuda):
    """Statuses of runs stopped early, forced from the host: an IPM
    iteration cap gives "max_iter"; a CG cap of one iteration per Newton
    system still ends (the reference ignores the CG status, fft_kkt.jl:353)
    with a finite objective; a NaN sample in z0 gives "failure" after the
    first residual check instead of iterating to the cap."""
    import csipm.workloads as wl
    from csipm.fft_operator import FFTParameters
    from csipm.model import FFTNLPModel, ipm_solve
    print("failure statuses", flush=True)
    p = wl.batch_instance((63, 63), seed=71)
    for mixed in (False, True):
        nlp = FFTNLPModel(FFTParameters(2, p.dims, p.z0, p.lam, p.missing), mixed=mixed)
        st = ipm_solve(nlp, tol=1e-8, max_iter=2)
        assert st.status == "max_iter" and st.iterations == 2, (mixed, st.status)
        st = ipm_solve(nlp, tol=1e-8, max_iter=40, cg_itmax=1)
        assert st.status in ("max_iter", "solved", "failure") and st.iterations <= 40, mixed
        # at most one CG iteration per Newton system (predictor, corrector and,
        # mixed, the refinement correction)
        assert st.cg_iterations <= 3 * (st.iterations + 1), (mixed, st.cg_iterations)
        z = p.z0.copy()
        z[5] = np.nan
        bad = FFTNLPModel(FFTParameters(2, p.dims, z, p.lam, p.missing), mixed=mixed)
        st = ipm_solve(bad, tol=1e-8, max_iter=50)
        assert st.status == "failure" and st.iterations <= 1, (mixed, st.status, st.iterations)
