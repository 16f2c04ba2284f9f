"""Serving benchmark: an open-loop Poisson request stream, continuous batching
(mixed precision) against plain FP64 one-solve-at-a-time.

    python compressedsensingipm.jl_amd/csipm/serve_bench.py --requests 400 --rate 50

Requests follow the serving workload: Poisson arrivals at `rate` per second;
70% 1D (N log-uniform [1024, 65536]), 25% 2D (odd side [127, 511]), 5% 3D
(odd side [63, 191]); each generated from its own seed by the generator of
1D / 2D / 3D benchmark generators at its shape (generation happens before the clock starts).

* batched: csipm.batch.serve's loop -- a request's operator handle is made
  by a helper thread when it arrives; it is admitted once that is ready and
  its device footprint fits the budget, into a running batch engine (FP32
  operator and CG, FP64 control); one interior-point iteration of every
  admitted request per engine step; finished ones retire.
* plain: the same requests in arrival order, one csipm_solve at a time on
  FP64 handles, each started when it has arrived and the previous is done.

Latency = completion - arrival.  Prints one JSON line: solves/s over the
makespan, p50 / p99 latency of both arms, and the objective agreement of a
sample: batched vs FP64 (relative gap) and batched vs the same request solved
alone in the mixed form.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=400)
    ap.add_argument("--rate", type=float, default=50.0)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--pool-gb", type=float, default=16.0)
    ap.add_argument("--arms", default="batched,plain")
    ap.add_argument("--check", type=int, default=12, help="requests re-solved for agreement")
    ap.add_argument("--prep-workers", type=int, default=4,
                    help="threads making operator handles for arriving requests")
    args = ap.parse_args()

    import numpy as np
    import torch

    import csipm.workloads as wl
    from csipm.batch import BatchSolver, pool_bytes, slots_of, solve_batch
    from csipm.fft_operator import FFTOperator, FFTParameters
    from csipm.model import FFTNLPModel, ipm_solve

    dev = torch.device("cuda:0")
    arrivals, shapes, _ = wl.stream_requests(nreq=args.requests, rate=args.rate)
    t0 = time.perf_counter()
    probs = []
    for i, s in enumerate(shapes):
        probs.append(wl.stream_problem(s, seed=1_000_003 * 4 + 1 + i))
        if (i + 1) % 100 == 0:
            print(f"# generated {i + 1}", file=sys.stderr, flush=True)
    print(f"# generated {len(probs)} requests in {time.perf_counter() - t0:.1f} s "
          f"({sum(len(s) == 3 for s in shapes)} 3D)", file=sys.stderr, flush=True)
    z0s = [torch.from_numpy(p.z0).to(dev) for p in probs]
    out = {"benchmark": "serving: Poisson stream, continuous batching (mixed) vs plain FP64",
           "requests": len(probs), "rate_per_s": args.rate, "tol": args.tol,
           "mix": {"1d": sum(len(s) == 1 for s in shapes), "2d": sum(len(s) == 2 for s in shapes),
                   "3d": sum(len(s) == 3 for s in shapes)},
           "budget_gb": args.budget_gb}

    def pct(lat, q):
        return round(1e3 * float(np.percentile(lat, q)), 2)

    res_b = {}
    if "batched" in args.arms:
        from concurrent.futures import ThreadPoolExecutor
        pool_slots = int(args.pool_gb * 2**30) // pool_bytes(1, True)
        bs = BatchSolver(pool_slots, max_instances=1024, mixed=True)
        budget = args.budget_gb * 2**30
        waiting, live_bytes, ops = [], {}, {}
        # operator handles (tables, mask, plans) are made by helper threads as
        # requests arrive, beside the engine steps of the main thread
        prep = ThreadPoolExecutor(max_workers=args.prep_workers)

        def make_op(k):
            p = probs[k]
            with torch.cuda.device(dev):
                return FFTOperator(p.N, len(p.dims), p.dims, p.missing, mixed=True)

        futs = {}
        k_next = 0
        t_start = time.perf_counter()
        while k_next < len(probs) or waiting or bs.live:
            now = time.perf_counter() - t_start
            while k_next < len(probs) and arrivals[k_next] <= now:
                waiting.append(k_next)
                futs[k_next] = prep.submit(make_op, k_next)
                k_next += 1
            still = []
            for k in waiting:
                p = probs[k]
                if k not in ops:
                    if not futs[k].done():
                        still.append(k)
                        continue
                    ops[k] = futs.pop(k).result()
                op = ops[k]
                need = pool_bytes(slots_of(op), True) + 48 * p.N
                if sum(live_bytes.values()) + need > budget or \
                        bs.admit(op, z0s[k], p.lam, tag=k) is None:
                    still.append(k)
                    continue
                live_bytes[k] = need
            waiting = still
            if bs.live == 0:
                if k_next < len(probs) or waiting:
                    time.sleep(max(0.0002, min(0.002, arrivals[min(k_next, len(probs) - 1)] - now)))
                continue
            bs.step(args.tol)
            for i in bs.finished():
                r = bs.harvest(i)
                k = r.tag
                r.t_done = time.perf_counter() - t_start
                res_b[k] = r
                live_bytes.pop(k, None)
                ops.pop(k, None)
        span = time.perf_counter() - t_start
        engine_stats = bs.stats()
        bs.close()
        prep.shutdown()
        lat = [res_b[k].t_done - arrivals[k] for k in res_b]
        out["batched_mixed"] = {"solves_per_s": round(len(res_b) / span, 2),
                                "makespan_s": round(span, 2), "p50_ms": pct(lat, 50),
                                "p99_ms": pct(lat, 99),
                                "all_solved": all(r.status == "solved" for r in res_b.values()),
                                "not_solved": [[k, list(shapes[k]), r.status, r.iterations,
                                                r.cg_iterations]
                                               for k, r in sorted(res_b.items())
                                               if r.status != "solved"][:10],
                                "max_cg_iterations": max(r.cg_iterations
                                                         for r in res_b.values()),
                                "engine": {k: round(v, 3) for k, v in engine_stats.items()}}
        print(f"# batched: {span:.1f} s", file=sys.stderr, flush=True)

    res_p = {}
    if "plain" in args.arms:
        t_start = time.perf_counter()
        lat = []
        for k, p in enumerate(probs):
            now = time.perf_counter() - t_start
            if arrivals[k] > now:
                time.sleep(arrivals[k] - now)
            m = FFTNLPModel(FFTParameters(len(p.dims), p.dims, z0s[k], p.lam, p.missing),
                            dtype=torch.float64)
            st = ipm_solve(m, tol=args.tol)
            torch.cuda.synchronize()
            res_p[k] = st
            lat.append(time.perf_counter() - t_start - arrivals[k])
            del m
        span = time.perf_counter() - t_start
        out["plain_fp64"] = {"solves_per_s": round(len(res_p) / span, 2),
                             "makespan_s": round(span, 2), "p50_ms": pct(lat, 50),
                             "p99_ms": pct(lat, 99),
                             "all_solved": all(r.status == "solved" for r in res_p.values())}
        print(f"# plain: {span:.1f} s", file=sys.stderr, flush=True)

    if res_b and res_p:
        out["batched_over_plain_p99"] = round(out["plain_fp64"]["p99_ms"] /
                                              out["batched_mixed"]["p99_ms"], 3)
        gaps = [abs(res_b[k].objective - res_p[k].objective) / abs(res_p[k].objective)
                for k in res_b if k in res_p]
        out["max_obj_rel_gap_vs_fp64"] = max(gaps)
    if res_b and args.check > 0:
        # the same request alone in the mixed form (a batch of one)
        idx = np.linspace(0, len(probs) - 1, min(args.check, len(probs))).astype(int)
        diffs = []
        for k in idx:
            p = probs[k]
            op = FFTOperator(p.N, len(p.dims), p.dims, p.missing, mixed=True)
            alone = solve_batch([op], [z0s[k]], [p.lam], mixed=True, tol=args.tol)[0]
            diffs.append(abs(alone.objective - res_b[k].objective) / abs(alone.objective))
        out["alone_vs_mix_max_obj_rel_diff"] = max(diffs)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
