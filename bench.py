"""Benchmark: compressed-sensing IPM solves on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 2d255] [--batch B]

Workload 2d255 (default): B instances (default 256) of the 2D 255x255
generator (40 modes, r=3 discs on a 16-lattice, lambda 0.05, tol 1e-8,
instance seeds 1..B), already resident in HBM.  (1d1024: B 1D N=1024 instances; 3d255: one 3D
255^3 instance; uneven: the 384 uneven 1D/2D instances at tol 1e-6.)
A step solves all B:

* batched mixed (the measured form, `value`): one batch engine, all B
  admitted at once, every kernel launch covering all of them; FP32 operator
  and CG on the reduced Newton system, FP64 control and refinement;
* plain FP64 (the baseline form, reported beside): the same B instances
  solved one at a time by csipm_solve on FP64 handles (the reference's
  condensed-K CG in FP64).

`value` is solves/s of the batched mixed form over the K timed steps; the
plain FP64 rate, the ratio, per-solve p50/p99 latencies of both forms (a
batched solve completes with its batch) and the worst relative objective gap
of the mixed solutions against the FP64 ones are reported beside it.  Also:

* roofline: the dominant launch of the batch's CG operator apply (the
  reduced-system M^T M + D apply over all B instances), timed live with HIP
  events (csipm_batch_apply_timed) on the stream the engine launches on;
  algorithmic bytes of the launch / its average duration, against 8 TB/s.
* cpu_baseline: the FP64 CPU restatement of the reference path (oracle/,
  scipy pocketfft) solving instance 1, a bounded sample of at most ~20 s.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "compressedsensingipm.jl_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "1d1024": ("signal_1d_1024", 1e-8,
               "1D N=1024, 12 modes, gaps+5% dropout, lambda 0.1, tol 1e-8"),
    "2d255": ("punched_2d_255", 1e-8,
                "2D 255x255, 40 modes, r=3 discs on a 16-lattice, lambda 0.05, tol 1e-8"),
    "3d255": ("punched_3d", 1e-8,
                "3D 255^3, 200 modes, r=2 balls on an 8-lattice + 1% dropout, lambda 0.05, "
                "tol 1e-8"),
    "uneven": ("uneven_batch", 1e-6,
                "256 1D (N log-uniform [512, 65536]) + 128 2D (odd side [63, 511]) instances, "
                "10% dropout + gaps/discs, lambda 0.1, tol 1e-6"),
}


def make_problems(wl, workload, batch):
    if workload == "uneven":
        return wl.uneven_batch()
    if workload == "3d255":
        return [wl.punched_3d(instance=1)]
    gen = getattr(wl, WORKLOADS[workload][0])
    return [gen(instance=i + 1) for i in range(batch)]


def load_traffic(workload, kernel):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            data = json.load(f)
        return data.get(workload, {}).get(kernel)
    except (OSError, ValueError):
        return None


def cpu_baseline(problem, tol, budget_s=20.0, max_runs=5, applies_per_solve=None):
    import numpy as np
    from oracle import ipm
    from oracle import operator as op
    workers = max(1, min(16, os.cpu_count() or 1))
    b = op.m_perpt_z(problem.z0, problem.dims, workers=workers)
    mt = lambda v: op.mtm(v, problem.dims, problem.missing, workers=workers)   # noqa: E731
    if problem.N > 2_000_000 and applies_per_solve:
        # a full CPU solve takes minutes here: time operator applies (the
        # solve's dominant cost) and scale by the applies of one FP64 solve
        x = np.random.default_rng(0).standard_normal(problem.N)
        times = []
        t_start = time.perf_counter()
        while len(times) < 5 and (time.perf_counter() - t_start) < budget_s:
            t0 = time.perf_counter()
            mt(x)
            times.append(time.perf_counter() - t0)
        med = statistics.median(times)
        est = med * applies_per_solve
        return {"value": 1.0 / est, "unit": "solves/s", "cores": workers, "kind": "port",
                "sample": f"{len(times)} FP64 oracle M^T M applies of instance 1 (median "
                          f"{med:.3f} s, scipy pocketfft workers={workers}) x "
                          f"{applies_per_solve} applies per FP64 solve (the GPU FP64 solve's "
                          "CG + residual count); solve not run to the end"}, None
    times = []
    res = None
    t_start = time.perf_counter()
    while len(times) < max_runs and (time.perf_counter() - t_start) < budget_s:
        t0 = time.perf_counter()
        res = ipm.solve(b, problem.lam, mt, ipm.IPMOptions(tol=tol))
        times.append(time.perf_counter() - t0)
    f_ref = ipm.objective(res.beta, np.abs(res.beta), mt(res.beta), b, problem.lam)
    med = statistics.median(times)
    return {"value": 1.0 / med, "unit": "solves/s", "cores": workers, "kind": "port",
            "sample": f"{len(times)} FP64 oracle solve(s) of instance 1 "
                      f"(median {med:.3f} s, scipy pocketfft workers={workers}, "
                      f"{res.iterations} IPM iterations)"}, f_ref


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="2d255", choices=sorted(WORKLOADS))
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every "
              "instance fits one MI355X)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import csipm.workloads as wl
    from csipm.batch import BatchSolver, slots_of
    from csipm.fft_operator import FFTOperator, FFTParameters
    from csipm.model import FFTNLPModel, ipm_solve

    dev = torch.device("cuda:0")
    gen, tol, desc = WORKLOADS[args.workload]
    problems = make_problems(wl, args.workload, args.batch)
    B = len(problems)
    z0s = [torch.from_numpy(p.z0).to(dev) for p in problems]
    print(f"# {args.workload}: {B} instances, dims {problems[0].dims}", file=sys.stderr,
          flush=True)

    models = [FFTNLPModel(FFTParameters(len(p.dims), p.dims, z, p.lam, p.missing),
                          dtype=torch.float64) for p, z in zip(problems, z0s)]
    mops = [FFTOperator(p.N, len(p.dims), p.dims, p.missing, mixed=True) for p in problems]
    bs = BatchSolver(sum(slots_of(o) for o in mops), max_instances=B, mixed=True)

    def run_plain():
        out, lat = [], []
        for m in models:
            t0 = time.perf_counter()
            out.append(ipm_solve(m, tol=tol))
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t0)
        return out, lat

    def run_batched():
        t0 = time.perf_counter()
        for k, (o, p, z) in enumerate(zip(mops, problems, z0s)):
            if bs.admit(o, z, p.lam, tag=k) is None:
                raise RuntimeError("pool too small")
        while bs.step(tol) > 0:
            pass
        res = [None] * B
        for i in bs.finished():
            r = bs.harvest(i)
            res[r.tag] = r
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        return res, [t1 - t0] * B

    for _ in range(args.warmup):
        run_batched()
        run_plain()
    torch.cuda.synchronize()
    print("# warm", file=sys.stderr, flush=True)
    t_all = time.perf_counter()
    lat_b, res_b = [], None
    for _ in range(args.steps):
        res_b, lat = run_batched()
        lat_b += lat
    torch.cuda.synchronize()
    total_b = time.perf_counter() - t_all
    print(f"# batched mixed: {total_b:.3f} s", file=sys.stderr, flush=True)
    t_all = time.perf_counter()
    lat_p, res_p = [], None
    for _ in range(args.steps):
        res_p, lat = run_plain()
        lat_p += lat
    torch.cuda.synchronize()
    total_p = time.perf_counter() - t_all
    print(f"# plain fp64: {total_p:.3f} s", file=sys.stderr, flush=True)

    gaps = [abs(rb.objective - rp.objective) / abs(rp.objective) for rb, rp in zip(res_b, res_p)]
    beta_err = [float(np.linalg.norm(rb.beta.cpu().numpy() - rp.solution.cpu().numpy()) /
                      np.linalg.norm(rp.solution.cpu().numpy()))
                for rb, rp in zip(res_b[:8], res_p[:8])]
    statuses = [r.status for r in res_b] + [r.status for r in res_p]

    # roofline: the dominant launch of the batch's CG operator apply
    for k, (o, p, z) in enumerate(zip(mops, problems, z0s)):
        bs.admit(o, z, p.lam, tag=k)
    bs.apply_timed(5)
    launches = bs.apply_timed(50)
    name, lms, lbytes = max(launches, key=lambda t: t[1])
    achieved = lbytes / (lms * 1e-3) / 1e9
    apply_ms = sum(t[1] for t in launches)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": load_traffic(f"{args.workload}-batch{B}-mixed", name),
                "kernel": name, "kernel_ms": round(lms, 5), "kernel_bytes": int(lbytes),
                "apply_ms": round(apply_ms, 5),
                "apply_GBps": round(sum(t[2] for t in launches) / (apply_ms * 1e-3) / 1e9, 1),
                "launches": {t[0]: {"ms": round(t[1], 5), "bytes": int(t[2])} for t in launches}}

    n_solves = B * args.steps
    value = n_solves / total_b
    plain = n_solves / total_p
    p0 = problems[0]
    out = {
        "metric": "solves/s at <=1e-6 obj gap, batched FP32-op vs plain FP64; p50/p99 latency",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total_b / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-op/f64-control",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {desc}; a step = {B} instance(s) solved "
                               "batched (FP32 operator/CG, FP64 control) in one engine",
                   "dims": list(p0.dims) if B == 1 or args.workload != "uneven" else "mixed",
                   "N": p0.N if args.workload != "uneven" else int(sum(p.N for p in problems)),
                   "batch": B, "lambda": p0.lam, "tol": tol},
        "plain_fp64": {"value": round(plain, 3), "unit": "solves/s",
                       "p50_ms": round(1e3 * statistics.median(lat_p), 3),
                       "p99_ms": round(1e3 * float(np.percentile(lat_p, 99)), 3),
                       "ipm_iterations": res_p[0].iterations,
                       "cg_iterations": res_p[0].cg_iterations},
        "batched_over_plain": round(value / plain, 3),
        "p50_ms": round(1e3 * statistics.median(lat_b), 3),
        "p99_ms": round(1e3 * float(np.percentile(lat_b, 99)), 3),
        "ipm_iterations": res_b[0].iterations,
        "cg_iterations": res_b[0].cg_iterations,
        "max_obj_rel_gap_vs_fp64": max(gaps),
        "beta_rel_l2_vs_fp64": max(beta_err),
        "all_solved": all(s == "solved" for s in statuses),
        "roofline": roofline,
    }
    bs.close()
    if not args.no_cpu_baseline:
        applies = res_p[0].cg_iterations + 3 * res_p[0].iterations
        cb, f_ref = cpu_baseline(problems[0], tol, applies_per_solve=applies)
        out["cpu_baseline"] = cb
        if f_ref is not None:
            out["obj_rel_gap_fp64_vs_cpu"] = abs(res_p[0].objective - f_ref) / abs(f_ref)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
