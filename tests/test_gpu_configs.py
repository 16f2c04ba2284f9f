"""GPU parity on the benchmark configurations at their full sizes.

* 2D 255x255 (FP64 on the GPU, tol 1e-8): FP64 and mixed solves against the
  FP64 oracle solved with the same settings (the reference's CPU path, same
  tolerance and CG rule): objective within 1e-6, the same support above
  1e-6 max|beta| (every flip listed, none exempt); objective within 1e-6 of
  the oracle solved to 1e-10.
* 3D 255^3 mixed precision (tol 1e-8): FP64 and FP32 operator applies
  against the oracle at full size; the mixed solve against the FP64 GPU
  solve (objective within 1e-6, beta relative l2, support) and the LASSO
  certificate of both at full size: the duality gap of the residual-scaled
  dual point, an upper bound of P(beta) - P* whatever the algorithm, within
  twice the total complementarity 2N mu the stopping rule leaves (at
  N = 16.6 M and tol 1e-8 that is about 1e-5 of P).
* The 384 uneven instances (256 1D + 128 2D, FP32 operator / FP64 control,
  tol 1e-6) in one batch: every instance solved; each against the plain
  FP64 solve of the same instance (objective within 1e-6, the metric's
  gap); each certified (the duality gap of the residual-scaled dual point,
  an upper bound of P(beta) - P*, within 1e-2 P at this tolerance); a
  16-instance sample spanning direct, Bluestein,
  four-step and four-step Bluestein lengths against the oracle at the same
  tolerance (objective within 1e-6), and re-solved batched at tol 1e-8 for
  the support rule against the oracle at 1e-8 (a flip on which the batched
  solution agrees with the oracle solved to 1e-10 is the 1e-8 oracle's own
  unconverged entry; it is printed, every other flip fails).
* The serving stream's first 200 requests (including its 3D requests)
  through csipm.batch.serve at their arrival times: every request solved,
  each result the same as the request solved alone (objective 1e-9, beta
  1e-6, the same support) and within 1e-6 (objective) of the plain FP64
  solve of the request, the metric's gap.

At tol 1e-6 an interior-point iterate still carries off-support entries of
up to ~1e-4 max|beta| (the oracle itself, at 1e-6 against 1e-10, flips
hundreds of entries at 1e-6 max|beta|), so the support rule is applied
where the solves are run to 1e-8.
"""
import numpy as np
import pytest
import torch

import _checks as chk
from oracle import operator as op

pytestmark = pytest.mark.gpu


def _op(p, mixed=False, dtype=torch.float64):
    from csipm.fft_operator import FFTOperator
    return FFTOperator(p.N, len(p.dims), p.dims, p.missing, mixed=mixed, dtype=dtype)


def _solve(p, mixed, tol, **kw):
    from csipm.fft_operator import FFTParameters
    from csipm.model import FFTNLPModel, ipm_solve
    params = FFTParameters(len(p.dims), p.dims, p.z0, p.lam, p.missing)
    return ipm_solve(FFTNLPModel(params, mixed=mixed), tol=tol, **kw)


def test_2d255_solves_vs_oracle(cuda):
    import csipm.workloads as wl
    print("2d255 solves vs oracle", flush=True)
    p = wl.punched_2d_255(instance=1)
    ref8, f8 = chk.oracle_solve(p, 1e-8)
    ref10, f10 = chk.oracle_solve(p, 1e-10, cg_rtol=1e-13)
    for mixed in (False, True):
        st = _solve(p, mixed, 1e-8)
        assert st.status == "solved"
        beta = st.solution.cpu().numpy()
        f = chk.objective(p, beta)
        name = "mixed" if mixed else "fp64"
        print(f"  {name}: gap vs oracle(1e-8) {abs(f - f8) / abs(f8):.2e}, vs oracle(1e-10) "
              f"{abs(f - f10) / abs(f10):.2e}", flush=True)
        assert abs(f - f8) <= 1e-6 * abs(f8), name
        assert abs(f - f10) <= 1e-6 * abs(f10), name
        nsup = chk.assert_same_support(beta, ref8.beta, f"2d255 {name}")
        assert nsup > 0


def test_3d255_operator_applies(cuda):
    import csipm.workloads as wl
    from csipm.fft_operator import M_perpt_M_perp_vec
    print("3d255 operator applies", flush=True)
    p = wl.punched_3d(instance=1)
    x = np.random.default_rng(3).standard_normal(p.N)
    ref = op.mtm(x, p.dims, p.missing, workers=8)
    for dtype, bound in ((torch.float64, 1e-12), (torch.float32, 2e-5)):
        o = _op(p, dtype=dtype)
        y = M_perpt_M_perp_vec(o, torch.from_numpy(x).to(cuda, dtype)).double().cpu().numpy()
        err = float(np.abs(y - ref).max() / np.abs(ref).max())
        print(f"  {dtype}: rel err {err:.2e}", flush=True)
        assert err <= bound, (dtype, err)
        del o


def test_3d255_mixed_solve(cuda):
    import csipm.workloads as wl
    print("3d255 mixed solve", flush=True)
    p = wl.punched_3d(instance=1)
    s64 = _solve(p, False, 1e-8)
    smx = _solve(p, True, 1e-8)
    assert s64.status == "solved" and smx.status == "solved"
    b64 = s64.solution.cpu().numpy()
    bmx = smx.solution.cpu().numpy()
    gap = abs(smx.objective - s64.objective) / abs(s64.objective)
    rel = float(np.linalg.norm(bmx - b64) / np.linalg.norm(b64))
    print(f"  objective gap {gap:.2e}, beta rel l2 {rel:.2e}", flush=True)
    assert gap <= 1e-6 and rel <= 1e-6
    chk.assert_same_support(bmx, b64, "3d255 mixed vs fp64")
    for name, st in (("fp64", s64), ("mixed", smx)):
        fp, dgap, primal = chk.certificate(p, st.solution.cpu().numpy())
        # the stopping rule leaves a total complementarity 2N mu (mu <=
        # tol (1 + |f| / N)); P(beta) - P* <= dgap holds whatever the algorithm
        comp = 2 * p.N * st.mu
        print(f"  {name}: certificate gap {dgap / primal:.2e} of P (2N mu {comp / primal:.2e} "
              f"of P), fixed-point {fp:.2e}", flush=True)
        assert dgap <= 2.0 * comp + 1e-6 * abs(primal), name


def _uneven_sample(shapes):
    """16 instance indices: the smallest of each length class (1D direct,
    1D Bluestein, four-step, four-step Bluestein, 2D direct, 2D Bluestein)."""
    def smooth(n):
        for r in (2, 3, 5, 7, 11, 13, 17):
            while n % r == 0:
                n //= r
        return n == 1
    cls = {}
    for i, s in enumerate(shapes):
        if len(s) == 1:
            k = ("1d", s[0] > 4096, smooth(s[0]))
        else:
            k = ("2d", smooth(s[0]))
        cls.setdefault(k, []).append(i)
    take = []
    for k, ids in sorted(cls.items(), key=lambda t: str(t[0])):
        ids = sorted(ids, key=lambda i: int(np.prod(shapes[i])))
        take += ids[:3 if k[0] == "1d" else 2]
    return take


def test_uneven_batch_384(cuda):
    import csipm.workloads as wl
    from csipm.batch import solve_batch
    print("uneven batch of 384", flush=True)
    probs = wl.uneven_batch()
    assert len(probs) == 384
    ops = [_op(p, mixed=True) for p in probs]
    res = solve_batch(ops, [p.z0 for p in probs], [p.lam for p in probs], mixed=True, tol=1e-6)
    del ops
    assert all(r.status == "solved" for r in res), [r.status for r in res]
    worst_gap, worst_cert = 0.0, 0.0
    for k, (p, r) in enumerate(zip(probs, res)):
        s64 = _solve(p, False, 1e-6)
        assert s64.status == "solved", k
        gap = abs(r.objective - s64.objective) / abs(s64.objective)
        worst_gap = max(worst_gap, gap)
        assert gap <= 1e-6, (k, p.dims, gap)
        fp, dgap, primal = chk.certificate(p, r.beta.cpu().numpy(), workers=4)
        worst_cert = max(worst_cert, dgap / abs(primal))
        # at tol 1e-6 the residual-scaled dual point is up to a few 1e-3 P from
        # optimal: the certificate is a sanity bound here (the parity is the
        # objective against the FP64 solve and the oracle)
        assert dgap <= 1e-2 * abs(primal), (k, p.dims, dgap / primal)
    print(f"  worst gap vs plain fp64 {worst_gap:.2e}, worst certificate {worst_cert:.2e}",
          flush=True)
    shapes = [p.dims for p in probs]
    sample = _uneven_sample(shapes)
    assert len(sample) >= 12
    for k in sample:
        p = probs[k]
        _, f6 = chk.oracle_solve(p, 1e-6)
        f = chk.objective(p, res[k].beta.cpu().numpy())
        assert abs(f - f6) <= 1e-6 * abs(f6), (k, p.dims, abs(f - f6) / abs(f6))
    # the sample solved batched to 1e-8: support rule against the oracle at 1e-8
    sp = [probs[k] for k in sample]
    res8 = solve_batch([_op(p, mixed=True) for p in sp], [p.z0 for p in sp],
                       [p.lam for p in sp], mixed=True, tol=1e-8)
    for p, r in zip(sp, res8):
        assert r.status == "solved", p.dims
        ref8, f8 = chk.oracle_solve(p, 1e-8)
        ref10, _ = chk.oracle_solve(p, 1e-10, cg_rtol=1e-13)
        beta = r.beta.cpu().numpy()
        assert abs(chk.objective(p, beta) - f8) <= 1e-6 * abs(f8), p.dims
        _, excused = chk.assert_same_support(beta, ref8.beta, f"uneven {p.dims} tol 1e-8",
                                             tighter=ref10.beta)
        if excused:
            print(f"  {p.dims}: oracle(1e-8) entries {excused} unconverged, beta agrees "
                  "with oracle(1e-10)", flush=True)


def test_stream_prefix_served(cuda):
    import csipm.workloads as wl
    from csipm.batch import Request, serve, solve_batch
    print("stream prefix through serve()", flush=True)
    n = 200
    arrivals, shapes, _ = wl.stream_requests()
    probs = [wl.stream_problem(shapes[k], seed=1_000_003 * 4 + 1 + k) for k in range(n)]
    n3 = sum(len(p.dims) == 3 for p in probs)
    assert n3 > 0
    z0s = [torch.from_numpy(p.z0).to(cuda) for p in probs]

    class Q:
        def __init__(self, p, z):
            self.p, self.dims, self.z0, self.lam, self.missing = p, p.dims, z, p.lam, p.missing

    reqs = [Request(float(arrivals[k]), k, Q(probs[k], z0s[k])) for k in range(n)]
    out = serve(reqs, lambda q: _op(q.p, mixed=True), mixed=True, budget_bytes=64 << 30,
                tol=1e-6, prep_workers=4)
    assert sorted(r.tag.seq for r in out) == list(range(n))
    print(f"  {n} requests ({n3} 3D) served", flush=True)
    worst, worst64 = 0.0, 0.0
    for r in out:
        k = r.tag.seq
        p = probs[k]
        assert r.status == "solved", (k, p.dims, r.status, r.error)
        alone = solve_batch([_op(p, mixed=True)], [z0s[k]], [p.lam], mixed=True, tol=1e-6)[0]
        gap = abs(r.objective - alone.objective) / abs(alone.objective)
        worst = max(worst, gap)
        assert gap <= 1e-9, (k, p.dims, gap)
        x = r.beta.cpu().numpy()
        y = alone.beta.cpu().numpy()
        assert np.linalg.norm(x - y) <= 1e-6 * np.linalg.norm(y), (k, p.dims)
        chk.assert_same_support(x, y, f"stream request {k} {p.dims}")
        # the metric's gap: the served mixed-precision result against the
        # plain FP64 solve of the same request at the same tolerance
        s64 = _solve(p, False, 1e-6)
        assert s64.status == "solved", (k, p.dims)
        g64 = abs(r.objective - s64.objective) / abs(s64.objective)
        worst64 = max(worst64, g64)
        assert g64 <= 1e-6, (k, p.dims, g64)
    print(f"  worst mix-vs-alone objective difference {worst:.2e}, worst gap vs plain FP64 "
          f"{worst64:.2e}", flush=True)
