"""Benchmark: compressed-sensing IPM solves on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload stream]

Workload `stream` (default): the serving stream of 10,000 requests, each from
its own seed: 70% 1D (N log-uniform [1024, 65536], 12 + N/128 modes, 3 gaps
of 32 + 5% dropout, lambda 0.1), 25% 2D (odd side [127, 511], 40 modes, r=3
discs on a 16-lattice, lambda 0.05), 5% 3D (odd side [63, 191], 200 modes,
r=2 balls on an 8-lattice + 1% dropout, lambda 0.05), Poisson arrivals at
50/s, IPM tol 1e-6, device budget 200 GB.  A step is 500 consecutive requests
of the stream: K steps time the first 500*K requests (K = 20: the whole
stream); the W warm-up steps run the first 500*W requests untimed.  The
observed samples (z0) are resident in HBM before the clock starts, and so is
the server's device pool (150 GB of the 200 GB budget, allocated once as a
serving process does at start-up); every request's operator handle (plans,
sampling mask) is made inside the timed region, as a server must when the
request arrives.

* value: solves/s of the batched mixed-precision server (csipm.batch.serve:
  continuous batching, FP32 operator and CG, FP64 control and refinement;
  three engines, one each for the 3D, 2D and 1D requests, with half, a
  quarter and a quarter of the pool, whose turns run in parallel threads)
  at saturation -- all 500*K requests offered at t = 0.
* plain_fp64: the same requests solved one at a time by csipm_solve on FP64
  handles (FP64 condensed-K PCG, relative 1e-10), at saturation, and
  under the reference's PCG rule (atol 1e-12, rtol 0, src/fft_kkt.jl:339) on
  a prefix.
* latency: p50 / p99 of completion - arrival for both forms on the open-loop
  Poisson stream at 50/s over its first --latency-requests requests.
* roofline: the dominant launch of the batched CG operator apply over the
  stream's first 512 requests admitted together, HIP events on the engine's
  stream (csipm_batch_apply_timed); `--roofline-only` runs only that part
  (the command profiled with rocprofv3); `--batched-only` times the
  batched server alone (the stream's kernel-time profile).
* cpu_baseline: the FP64 CPU restatement of the reference path (oracle/,
  scipy pocketfft) on the host cores: the 1D N=1024 configuration and a
  fixed sample of stream requests.

Other workloads (--workload 2d255 | 1d1024 | 3d255 | uneven) time fixed
batches: a step solves --batch instances (2d255, 1d1024), one 255^3 instance
or the 384 uneven instances, batched-mixed vs plain FP64.  --workload
sweep3d is the reference's CPU-vs-GPU size sweep (benchmarks/cpu_vs_gpu.jl:
cubes 8^3 ... 560^3 of test/ipm_example_3D.jl's problem): per size the GPU
FP64 and mixed solve times, the batched FP32 apply's roofline, and the CPU
FP64 port's apply (and solve, up to 32^3).
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
METRIC = "solves/s at <=1e-6 obj gap, batched FP32-op vs plain FP64; p50/p99 latency"
STREAM_REQUESTS = 10_000
STREAM_RATE = 50.0
STREAM_TOL = 1e-6
PER_STEP = 500
BUDGET_BYTES = 200 << 30

WORKLOADS = {
    "stream": ("stream_problem", STREAM_TOL,
               "serving stream: 10,000 requests (70% 1D N log-uniform [1024, 65536], 25% 2D "
               "odd side [127, 511], 5% 3D odd side [63, 191]), Poisson 50/s, tol 1e-6, "
               "200 GB budget"),
    "1d1024": ("signal_1d_1024", 1e-8,
               "1D N=1024, 12 modes, gaps+5% dropout, lambda 0.1, tol 1e-8"),
    "2d255": ("punched_2d_255", 1e-8,
              "2D 255x255, 40 modes, r=3 discs on a 16-lattice, lambda 0.05, tol 1e-8"),
    "3d255": ("punched_3d", 1e-8,
              "3D 255^3, 200 modes, r=2 balls on an 8-lattice + 1% dropout, lambda 0.05, "
              "tol 1e-8"),
    "sweep3d": ("reference_example", 1e-8,
                "3D cubes 8^3 ... 560^3 (benchmarks/cpu_vs_gpu.jl), the problem of "
                "test/ipm_example_3D.jl (noise, punched holes, lambda 5), tol 1e-8"),
    "uneven": ("uneven_batch", 1e-6,
               "256 1D (N log-uniform [512, 65536]) + 128 2D (odd side [63, 511]) instances, "
               "10% dropout + gaps/discs, lambda 0.1, tol 1e-6"),
}


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def make_problems(wl, workload, batch):
    if workload == "uneven":
        return wl.uneven_batch()
    if workload == "3d255":
        return [wl.punched_3d(instance=1)]
    gen = getattr(wl, WORKLOADS[workload][0])
    return [gen(instance=i + 1) for i in range(batch)]


# ---------------------------------------------------------------------------
# stream generation (host, before the clock; one process per host core share)
# ---------------------------------------------------------------------------
def _gen_one(arg):
    import csipm.workloads as wl
    k, shape = arg
    p = wl.stream_problem(shape, seed=1_000_003 * 4 + 1 + k)
    return k, p.dims, p.z0, p.lam, p.missing


def stream_problems(n, workers):
    """The first n requests of the 10,000-request stream (arrival times,
    Problems)."""
    import multiprocessing as mp
    import csipm.workloads as wl
    arrivals, shapes, _ = wl.stream_requests(nreq=max(n, STREAM_REQUESTS), rate=STREAM_RATE)
    out = [None] * n
    t0 = time.perf_counter()
    # large requests first so that the pool stays busy to the end
    order = sorted(range(n), key=lambda k: -int(__import__("numpy").prod(shapes[k])))
    with mp.get_context("fork").Pool(workers) as pool:
        for i, (k, dims, z0, lam, miss) in enumerate(
                pool.imap_unordered(_gen_one, [(k, shapes[k]) for k in order], chunksize=4)):
            out[k] = wl.Problem(tuple(dims), z0, lam, miss, "x".join(map(str, dims)))
            if (i + 1) % 2000 == 0:
                log(f"generated {i + 1}/{n} requests ({time.perf_counter() - t0:.1f} s)")
    return arrivals[:n], out


def host_cpu():
    """(model name, physical cores, logical CPUs, CPUs this process may use)."""
    model, cores = "", set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if phys is not None and core is not None:
                        cores.add((phys, core))
                    phys = core = None
                    continue
                k, v = (t.strip() for t in line.split(":", 1))
                if k == "model name" and not model:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
        if phys is not None and core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    logical = os.cpu_count() or 1
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = logical
    return model, len(cores) or logical, logical, usable


def cpu_threads():
    """Threads of the CPU baseline: every CPU this process may run on
    (sched_getaffinity), bounded by the host's CPU share for this GPU where
    the environment states one (OMP_NUM_THREADS)."""
    usable = host_cpu()[3]
    try:
        share = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        share = 0
    return max(1, min(usable, share) if share > 0 else usable)


def cpu_solve_time(problem, tol, workers, runs, warm=1, budget_s=None):
    """Median seconds of oracle FP64 solves (construction of b = M^T z0 and
    the solve; signal, mask and FFT plans excluded), after `warm` untimed
    runs; returns (median, runs done, result of the last run)."""
    from oracle import ipm
    from oracle import operator as op
    mt = lambda v: op.mtm(v, problem.dims, problem.missing, workers=workers)   # noqa: E731
    times, res = [], None
    t_start = time.perf_counter()
    for r in range(warm + runs):
        t0 = time.perf_counter()
        b = op.m_perpt_z(problem.z0, problem.dims, workers=workers)
        res = ipm.solve(b, problem.lam, mt, ipm.IPMOptions(tol=tol))
        if r >= warm:
            times.append(time.perf_counter() - t0)
        if budget_s is not None and time.perf_counter() - t_start > budget_s and times:
            break
    return statistics.median(times), len(times), res


def cpu_baseline_stream(probs, shapes_frac, threads):
    """CPU restatement timed on the host: the 1D N=1024 configuration at 1
    and `threads` threads (median of 5 after 1 warm-up), and a fixed sample
    of stream requests at `threads` threads (the first 4 1D, 2 2D and the
    smallest 3D request of the timed set, one run each) whose per-type mean
    times, weighted by the stream's type shares, give the stream's solves/s."""
    import csipm.workloads as wl
    model, phys, logical, usable = host_cpu()
    p0 = wl.signal_1d_1024(instance=1)
    t1, n1, _ = cpu_solve_time(p0, 1e-8, 1, runs=5)
    tT, nT, _ = cpu_solve_time(p0, 1e-8, threads, runs=5)
    log(f"cpu 1d1024: {1 / t1:.1f} solves/s (1 thread), {1 / tT:.1f} ({threads} threads)")
    by = {1: [], 2: [], 3: []}
    for k, p in enumerate(probs):
        d = len(p.dims)
        if d == 3:
            continue
        if len(by[d]) < (4 if d == 1 else 2):
            by[d].append(p)
    three = [p for p in probs if len(p.dims) == 3]
    if three:
        by[3].append(min(three, key=lambda p: p.N))
    mean_t = {}
    for d, ps in by.items():
        if not ps:
            continue
        ts = [cpu_solve_time(p, STREAM_TOL, threads, runs=1, warm=0)[0] for p in ps]
        mean_t[d] = sum(ts) / len(ts)
        log(f"cpu stream {d}D sample: {[p.name for p in ps]} mean {mean_t[d]:.2f} s")
    t_mix = sum(shapes_frac[d] * mean_t[d] for d in mean_t) / \
        max(1e-12, sum(shapes_frac[d] for d in mean_t))
    sample = ("FP64 oracle (numpy + scipy pocketfft) solves at tol 1e-6 of the stream's first "
              f"{len(by[1])} 1D, {len(by[2])} 2D and smallest 3D request ("
              + ", ".join(p.name for d in (1, 2, 3) for p in by[d]) +
              "), one run each; per-type mean time weighted by the stream's 70/25/5 mix "
              "(the 3D mean rests on its smallest request, so this rate is an upper bound)")
    return {"value": round(1.0 / t_mix, 4), "unit": "solves/s", "cores": threads,
            "kind": "port", "sample": sample,
            "host_cpu": model, "host_physical_cores": phys, "host_logical_cpus": logical,
            "cpus_usable": usable,
            "config_1d1024": {"tol": 1e-8, "threads_1_solves_per_s": round(1.0 / t1, 3),
                              f"threads_{threads}_solves_per_s": round(1.0 / tT, 3),
                              "runs": [n1, nT], "statistic": "median after 1 warm-up"},
            "per_type_mean_s": {f"{d}d": round(t, 4) for d, t in mean_t.items()}}


# ---------------------------------------------------------------------------
# the stream workload
# ---------------------------------------------------------------------------
def roofline_of(bs):
    bs.apply_timed(5)
    launches = bs.apply_timed(50)
    name, lms, lbytes = max(launches, key=lambda t: t[1])
    achieved = lbytes / (lms * 1e-3) / 1e9
    apply_ms = sum(t[1] for t in launches)
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "kernel": name,
            "kernel_ms": round(lms, 5), "kernel_bytes": int(lbytes),
            "apply_ms": round(apply_ms, 5),
            "apply_GBps": round(sum(t[2] for t in launches) / (apply_ms * 1e-3) / 1e9, 1),
            "launches": {t[0]: {"ms": round(t[1], 5), "bytes": int(t[2])} for t in launches}}


def load_traffic(key, kernel):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            v = json.load(f).get(key, {}).get(kernel)
        return v, f"profiles/traffic.json[{key}][{kernel}]" if v is not None else None
    except (OSError, ValueError):
        return None, None


def stream_roofline(probs, z0s, dev):
    """Dominant launch of the batched CG apply over the first 512 requests
    (those that fit one 64 GB pool), admitted together."""
    from csipm.batch import BatchSolver, footprint
    from csipm.fft_operator import FFTOperator
    sel, slots = [], 0
    for k, p in enumerate(probs[:512]):
        s, _ = footprint(p.dims, len(p.missing), True)
        sel.append(k)
        slots += s
    bs = BatchSolver(slots, max_instances=len(sel), mixed=True)
    ops = [FFTOperator(probs[k].N, len(probs[k].dims), probs[k].dims, probs[k].missing,
                       mixed=True) for k in sel]
    for op, k in zip(ops, sel):
        if bs.admit(op, z0s[k], probs[k].lam, tag=k) is None:
            raise RuntimeError("roofline pool too small")
    roof = roofline_of(bs)
    roof["instances"] = len(sel)
    bs.close()
    del ops
    traffic, src = load_traffic("stream-first512-mixed", roof["kernel"])
    roof["traffic"] = traffic
    roof["traffic_source"] = src
    return roof


def run_stream(args):
    import numpy as np
    import torch

    import csipm.workloads as wl
    from csipm.batch import Request, serve
    from csipm.fft_operator import FFTOperator, FFTParameters
    from csipm.model import FFTNLPModel, ipm_solve

    dev = torch.device("cuda:0")
    n_timed = PER_STEP * args.steps
    n_warm = PER_STEP * args.warmup
    n_gen = max(n_timed, n_warm, 512 if args.roofline_only else 0)
    if args.roofline_only:
        n_gen = 512
    gen_workers = max(1, min(16, host_cpu()[3]))
    t0 = time.perf_counter()
    arrivals, probs = stream_problems(n_gen, gen_workers)
    log(f"generated {n_gen} requests in {time.perf_counter() - t0:.1f} s "
        f"({sum(len(p.dims) == 3 for p in probs)} 3D)")
    z0s = [torch.from_numpy(p.z0).to(dev) for p in probs]
    torch.cuda.synchronize()

    if args.roofline_only:
        roof = stream_roofline(probs, z0s, dev)
        print(json.dumps({"roofline": roof}), flush=True)
        return 0

    def make_mixed(p):
        return FFTOperator(p.N, len(p.dims), p.dims, p.missing, mixed=True)

    # the server's pool (150 GB of the 200 GB budget), made once at start-up
    # like a serving process's, outside the timed region: three engines
    # (lanes), the 3D, 2D and 1D requests', with half, a quarter and a
    # quarter of the pool
    from csipm.batch import BatchSolver, pool_bytes
    pool_slots = int(BUDGET_BYTES * 0.75) // pool_bytes(1, True)
    lanes = [(BatchSolver(pool_slots // 2, max_instances=256, mixed=True),
              lambda q: len(q.dims) == 3),
             (BatchSolver(pool_slots // 4, max_instances=512, mixed=True),
              lambda q: len(q.dims) == 2),
             (BatchSolver(pool_slots - pool_slots // 2 - pool_slots // 4,
                          max_instances=args.max_instances, mixed=True), None)]

    def batched(idx, arr):
        reqs = [Request(float(arr[j]), k, _Req(probs[k], z0s[k])) for j, k in enumerate(idx)]
        t_s = time.perf_counter()
        done = [0]

        def progress(r):
            done[0] += 1
            if done[0] % 2000 == 0:
                log(f"batched: {done[0]} requests, {time.perf_counter() - t_s:.1f} s")
        st = {}
        out = serve(reqs, lambda q: make_mixed(q.p), mixed=True, budget_bytes=BUDGET_BYTES,
                    max_instances=args.max_instances, tol=STREAM_TOL,
                    prep_workers=args.prep_workers, on_done=progress, stats=st, lanes=lanes)
        st["returned_s"] = time.perf_counter() - t_s
        torch.cuda.synchronize()
        st["synced_s"] = time.perf_counter() - t_s
        log("serve accounting: " + json.dumps({k: round(v, 3) for k, v in st.items()}))
        return out, time.perf_counter() - t_s

    def plain(idx, arr, cg_rule="relative", open_loop=False, cap_s=None):
        res, lat = {}, []
        t_s = time.perf_counter()
        for j, k in enumerate(idx):
            if open_loop:
                wait = arr[j] - (time.perf_counter() - t_s)
                if wait > 0:
                    time.sleep(wait)
            p = probs[k]
            m = FFTNLPModel(FFTParameters(len(p.dims), p.dims, z0s[k], p.lam, p.missing),
                            dtype=torch.float64)
            res[k] = ipm_solve(m, tol=STREAM_TOL, cg_rule=cg_rule)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t_s - (arr[j] if open_loop else 0.0))
            del m
            if (j + 1) % 1000 == 0:
                log(f"plain: {j + 1} requests, {time.perf_counter() - t_s:.1f} s")
            if cap_s is not None and time.perf_counter() - t_s > cap_s:
                break
        return res, lat, time.perf_counter() - t_s

    # warm-up: the first W steps' requests, batched; a few plain solves
    if n_warm:
        out_w, tw = batched(list(range(n_warm)), np.zeros(n_warm))
        log(f"warm-up batched: {len(out_w)} requests in {tw:.1f} s")
        plain(list(range(min(n_warm, 8))), None)
    idx = list(range(n_timed))
    zeros = np.zeros(n_timed)
    out_b, t_b = batched(idx, zeros)
    res_b = {r.tag.seq: r for r in out_b}
    value = len(out_b) / t_b
    solved_b = sum(r.status == "solved" for r in out_b)
    log(f"batched mixed at saturation: {n_timed} requests in {t_b:.2f} s ({value:.1f}/s, "
        f"{solved_b} solved)")
    if args.batched_only:
        print(json.dumps({"metric": METRIC, "value": round(value, 3), "unit": "solves/s",
                          "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": round(1e3 * t_b / args.steps, 3),
                          "higher_is_better": True, "requests": n_timed,
                          "solved": solved_b,
                          "note": "batched arm only (--batched-only): no plain FP64 arm, "
                                  "latency, roofline or CPU baseline"}), flush=True)
        return 0
    res_p, _, t_p = plain(idx, None, cap_s=args.plain_cap_s)
    n_p = len(res_p)
    plain_rate = n_p / t_p
    log(f"plain fp64 at saturation: {n_p} requests in {t_p:.2f} s ({plain_rate:.1f}/s)")
    # the reference's CG rule on a prefix
    n_ref = min(args.ref_rule_requests, n_timed)
    res_r, _, t_r = plain(list(range(n_ref)), None, cg_rule="reference")
    log(f"plain fp64, reference CG rule: {n_ref} requests in {t_r:.2f} s")
    rel_pref = sum(res_p[k].cg_iterations for k in range(n_ref) if k in res_p)
    ref_pref = sum(r.cg_iterations for r in res_r.values())
    # latency: open loop at 50/s over the first L requests
    L = min(args.latency_requests, n_timed)
    arr = np.asarray(arrivals[:L], dtype=float)
    out_l, _ = batched(list(range(L)), arr)
    lat_b = [r.t_done - arr[r.tag.seq] for r in out_l]
    res_lp, lat_p, _ = plain(list(range(L)), arr, open_loop=True)
    log(f"open-loop latency over {L} requests done")

    gap_keys = [k for k in res_p if k in res_b and res_b[k].status == "solved"]
    gaps = [abs(res_b[k].objective - res_p[k].objective) / abs(res_p[k].objective)
            for k in gap_keys]
    sample = [k for k in range(0, n_timed, max(1, n_timed // 64)) if k in res_p and k in res_b]
    beta_err = [float(torch.linalg.norm(res_b[k].beta - res_p[k].solution) /
                      torch.linalg.norm(res_p[k].solution)) for k in sample]
    mix = {d: sum(len(probs[k].dims) == d for k in idx) for d in (1, 2, 3)}

    roof = stream_roofline(probs, z0s, dev)

    def pct(v, q):
        return round(1e3 * float(np.percentile(v, q)), 2)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_b / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-op/f64-control",
        "data": "synthetic",
        "config": {"workload": "stream: " + WORKLOADS["stream"][2] +
                               f"; a step = {PER_STEP} consecutive requests, the first "
                               f"{n_timed} offered at t=0 (saturation) to the batched "
                               "mixed-precision server",
                   "requests": n_timed, "mix": {f"{d}d": mix[d] for d in mix},
                   "tol": STREAM_TOL, "budget_gb": BUDGET_BYTES >> 30,
                   "max_instances": args.max_instances,
                   "lanes": "3D engine (half the pool, 256 instances), 2D engine (a "
                            "quarter, 512), 1D engine (a quarter); turns in parallel "
                            "threads"},
        "all_solved": solved_b == n_timed and all(r.status == "solved" for r in res_p.values()),
        "not_solved": [[k, list(probs[k].dims), res_b[k].status] for k in res_b
                       if res_b[k].status != "solved"][:10],
        "plain_fp64": {"value": round(plain_rate, 3), "unit": "solves/s", "requests": n_p,
                       "cg_rule": "relative 1e-10",
                       "reference_cg_rule": {
                           "rule": "atol 1e-12, rtol 0 (src/fft_kkt.jl:339)",
                           "requests": n_ref, "solves_per_s": round(n_ref / t_r, 3),
                           "cg_iterations": ref_pref, "cg_iterations_relative_rule": rel_pref,
                           "all_solved": all(r.status == "solved" for r in res_r.values())}},
        "batched_over_plain": round(value / plain_rate, 3),
        "iterations_per_solve": {
            "batched_mixed": {"ipm": round(float(np.mean([r.iterations for r in out_b])), 2),
                              "cg": round(float(np.mean([r.cg_iterations for r in out_b])), 1)},
            "plain_fp64": {"ipm": round(float(np.mean([r.iterations for r in res_p.values()])), 2),
                           "cg": round(float(np.mean([r.cg_iterations
                                                      for r in res_p.values()])), 1)}},
        "latency_open_loop": {
            "rate_per_s": STREAM_RATE, "requests": L,
            "batched_mixed": {"p50_ms": pct(lat_b, 50), "p99_ms": pct(lat_b, 99)},
            "plain_fp64": {"p50_ms": pct(lat_p, 50), "p99_ms": pct(lat_p, 99)}},
        "p50_ms": pct(lat_b, 50),
        "p99_ms": pct(lat_b, 99),
        "max_obj_rel_gap_vs_fp64": max(gaps) if gaps else None,
        "obj_rel_gap_vs_fp64": {"compared": len(gaps),
                                "above_1e-6": int(sum(g > 1e-6 for g in gaps)),
                                "p50": float(np.percentile(gaps, 50)) if gaps else None,
                                "p99": float(np.percentile(gaps, 99)) if gaps else None,
                                "worst": sorted(([g, list(probs[k].dims)]
                                                 for k, g in zip(gap_keys, gaps)),
                                                reverse=True)[:5]},
        "beta_rel_l2_vs_fp64": max(beta_err) if beta_err else None,
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        frac = {1: 0.70, 2: 0.25, 3: 0.05}
        threads = cpu_threads()
        out["cpu_baseline"] = cpu_baseline_stream([probs[k] for k in idx], frac, threads)
    print(json.dumps(out), flush=True)
    return 0


class _Req:
    """A stream request as serve() sees it: the problem with its z0 already
    on the device."""

    def __init__(self, p, z0):
        self.p = p
        self.dims = p.dims
        self.z0 = z0
        self.lam = p.lam
        self.missing = p.missing


# ---------------------------------------------------------------------------
# fixed-batch workloads
# ---------------------------------------------------------------------------
def cpu_baseline_batch(problem, tol, threads, applies_per_solve=None):
    import numpy as np
    from oracle import ipm
    from oracle import operator as op
    if problem.N > 2_000_000 and applies_per_solve:
        x = np.random.default_rng(0).standard_normal(problem.N)
        times = []
        for _ in range(6):
            t0 = time.perf_counter()
            op.mtm(x, problem.dims, problem.missing, workers=threads)
            times.append(time.perf_counter() - t0)
        med = statistics.median(times[1:])
        return {"value": 1.0 / (med * applies_per_solve), "unit": "solves/s", "cores": threads,
                "kind": "port",
                "sample": f"5 FP64 oracle M^T M applies of instance 1 after 1 warm-up (median "
                          f"{med:.3f} s, scipy pocketfft workers={threads}) x "
                          f"{applies_per_solve} applies per FP64 solve; solve not run"}, None
    med, runs, res = cpu_solve_time(problem, tol, threads, runs=5, budget_s=30.0)
    mt = lambda v: op.mtm(v, problem.dims, problem.missing, workers=threads)   # noqa: E731
    b = op.m_perpt_z(problem.z0, problem.dims, workers=threads)
    f_ref = ipm.objective(res.beta, np.abs(res.beta), mt(res.beta), b, problem.lam)
    return {"value": 1.0 / med, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"median of {runs} FP64 oracle solve(s) of instance 1 after 1 warm-up "
                      f"({med:.3f} s, scipy pocketfft workers={threads}, "
                      f"{res.iterations} IPM iterations)"}, f_ref


SWEEP3D_SIDES = (8, 16, 32, 64, 96, 128, 192, 256, 384, 512, 560)


def run_sweep3d(args):
    """The reference's CPU-vs-GPU 3D size sweep (benchmarks/cpu_vs_gpu.jl:15-18):
    per cube, one FP64 and one mixed GPU solve on one handle (csipm_solve;
    sizes up to 64^3 after an untimed one), the batched FP32 CG apply's
    launches timed with HIP events (roofline of the dominant one and of the
    whole apply against 8 TB/s), and on the host the FP64 port's M^T M apply
    (up to 256^3) and solve (up to 32^3).  One JSON line, not the metric."""
    import numpy as np
    import torch
    import csipm.workloads as wl
    from csipm.batch import BatchSolver, slots_of
    from csipm.fft_operator import FFTOperator, FFTParameters
    from csipm.model import FFTNLPModel, ipm_solve
    dev = torch.device("cuda:0")
    tol = WORKLOADS["sweep3d"][1]
    threads = min(16, os.cpu_count() or 1)
    rows = []
    for n in SWEEP3D_SIDES:
        dims = (n, n, n)
        p = wl.reference_example(dims, check=False)
        z = torch.from_numpy(p.z0).to(dev)
        row = {"side": n, "N": n ** 3, "missing": int(len(p.missing))}
        for name, kw in (("fp64", {"dtype": torch.float64}), ("mixed", {"mixed": True})):
            # rdft = true, as the sweep runs the reference (cpu_vs_gpu.jl:9,27)
            m = FFTNLPModel(FFTParameters(3, dims, z, p.lam, p.missing), rdft=True, **kw)
            if n <= 64:
                ipm_solve(m, tol=tol)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = ipm_solve(m, tol=tol)
            torch.cuda.synchronize()
            row[name] = {"solve_s": round(time.perf_counter() - t0, 4), "status": r.status,
                         "ipm": r.iterations, "cg": r.cg_iterations,
                         "objective": r.objective}
            del m, r
        row["mixed_obj_rel_gap"] = abs(row["mixed"]["objective"] - row["fp64"]["objective"]) / \
            max(abs(row["fp64"]["objective"]), 1e-300)
        op = FFTOperator(p.N, 3, dims, p.missing, mixed=True)
        bs = BatchSolver(slots_of(op), max_instances=1, mixed=True)
        bs.admit(op, z, p.lam)
        bs.apply_timed(2)
        launches = bs.apply_timed(10)
        bs.close()
        op.close()
        dom = max(launches, key=lambda t: t[1])
        apply_ms = sum(t[1] for t in launches)
        row["apply"] = {"ms": round(apply_ms, 4),
                        "GBps": round(sum(t[2] for t in launches) / (apply_ms * 1e-3) / 1e9, 1),
                        "dominant": dom[0], "dominant_ms": round(dom[1], 4),
                        "dominant_frac": round(dom[2] / (dom[1] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        if not args.no_cpu_baseline and n <= 256:
            from oracle import operator as orc
            x = np.random.default_rng(0).standard_normal(p.N)
            orc.mtm(x, dims, p.missing, workers=threads)
            t0 = time.perf_counter()
            orc.mtm(x, dims, p.missing, workers=threads)
            row["cpu_port"] = {"apply_ms": round(1e3 * (time.perf_counter() - t0), 3),
                               "threads": threads}
            if n <= 32:
                med, runs, res = cpu_solve_time(p, tol, threads, runs=1, budget_s=20.0)
                row["cpu_port"]["solve_s"] = round(med, 4)
                row["cpu_port"]["ipm"] = res.iterations
        del z
        torch.cuda.empty_cache()
        log(f"sweep3d {n}^3: " + json.dumps(row))
        rows.append(row)
    print(json.dumps({"metric": "3D size sweep: solve seconds and apply roofline per cube side",
                      "unit": "s", "higher_is_better": False, "n_gpus": 1,
                      "workload": "sweep3d: " + WORKLOADS["sweep3d"][2],
                      "hbm_peak_GBps": HBM_PEAK_GBS, "sizes": rows}), flush=True)
    return 0


def run_batch_workload(args):
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
    log(f"{args.workload}: {B} instances, dims {problems[0].dims}")
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
        return res, [time.perf_counter() - t0] * B

    for _ in range(args.warmup):
        run_batched()
        run_plain()
    torch.cuda.synchronize()
    log("warm")
    t_all = time.perf_counter()
    lat_b, res_b = [], None
    for _ in range(args.steps):
        res_b, lat = run_batched()
        lat_b += lat
    torch.cuda.synchronize()
    total_b = time.perf_counter() - t_all
    log(f"batched mixed: {total_b:.3f} s")
    t_all = time.perf_counter()
    lat_p, res_p = [], None
    for _ in range(args.steps):
        res_p, lat = run_plain()
        lat_p += lat
    torch.cuda.synchronize()
    total_p = time.perf_counter() - t_all
    log(f"plain fp64: {total_p:.3f} s")
    gaps = [abs(rb.objective - rp.objective) / abs(rp.objective) for rb, rp in zip(res_b, res_p)]
    beta_err = [float(np.linalg.norm(rb.beta.cpu().numpy() - rp.solution.cpu().numpy()) /
                      np.linalg.norm(rp.solution.cpu().numpy()))
                for rb, rp in zip(res_b[:8], res_p[:8])]
    statuses = [r.status for r in res_b] + [r.status for r in res_p]
    for k, (o, p, z) in enumerate(zip(mops, problems, z0s)):
        bs.admit(o, z, p.lam, tag=k)
    roof = roofline_of(bs)
    roof["traffic"], roof["traffic_source"] = load_traffic(
        f"{args.workload}-batch{B}-mixed", roof["kernel"])
    n_solves = B * args.steps
    value = n_solves / total_b
    plain = n_solves / total_p
    p0 = problems[0]
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * total_b / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32-op/f64-control",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {desc}; a step = {B} instance(s) solved "
                               "batched (FP32 operator/CG, FP64 control) in one engine",
                   "dims": list(p0.dims) if args.workload != "uneven" else "mixed",
                   "batch": B, "tol": tol},
        "plain_fp64": {"value": round(plain, 3), "unit": "solves/s",
                       "p50_ms": round(1e3 * statistics.median(lat_p), 3),
                       "p99_ms": round(1e3 * float(np.percentile(lat_p, 99)), 3)},
        "batched_over_plain": round(value / plain, 3),
        "p50_ms": round(1e3 * statistics.median(lat_b), 3),
        "p99_ms": round(1e3 * float(np.percentile(lat_b, 99)), 3),
        "latency_note": "a batched solve completes with its batch",
        "max_obj_rel_gap_vs_fp64": max(gaps), "beta_rel_l2_vs_fp64": max(beta_err),
        "all_solved": all(s == "solved" for s in statuses), "roofline": roof,
    }
    bs.close()
    if not args.no_cpu_baseline:
        applies = res_p[0].cg_iterations + 3 * res_p[0].iterations
        threads = cpu_threads()
        cb, f_ref = cpu_baseline_batch(problems[0], tol, threads, applies_per_solve=applies)
        model, phys, logical, usable = host_cpu()
        cb.update({"host_cpu": model, "host_physical_cores": phys, "cpus_usable": usable})
        out["cpu_baseline"] = cb
        if f_ref is not None:
            out["obj_rel_gap_fp64_vs_cpu"] = abs(res_p[0].objective - f_ref) / abs(f_ref)
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="stream", choices=sorted(WORKLOADS))
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--max-instances", type=int, default=1024)
    ap.add_argument("--prep-workers", type=int, default=4)
    ap.add_argument("--latency-requests", type=int, default=500)
    ap.add_argument("--ref-rule-requests", type=int, default=100)
    ap.add_argument("--plain-cap-s", type=float, default=100.0,
                    help="stop the plain arm after this many seconds (reported)")
    ap.add_argument("--roofline-only", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="stream: time the batched server only (for profiling)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every "
              "instance fits one MI355X)", file=sys.stderr)
        return 2
    if args.workload == "stream":
        return run_stream(args)
    if args.workload == "sweep3d":
        return run_sweep3d(args)
    return run_batch_workload(args)


if __name__ == "__main__":
    sys.exit(main())
