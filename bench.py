"""Benchmark: full compressed-sensing IPM solves on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload config1]

A step is one complete solve (IPM with two PCG solves per iteration, every
operator apply through the fused HIP kernels) of one synthetic instance that
is already resident in HBM.  `value` is solves per second over the K timed
steps.  Also reported:

* roofline: the dominant pass of the fused condensed-KKT apply, timed live
  with HIP events (csipm_kkt_apply_timed) on the stream the library launches
  on; algorithmic bytes per launch / average duration, against 8 TB/s.
* cpu_baseline: the FP64 CPU restatement of the reference path (oracle/,
  scipy pocketfft) solving the same instance, run on this host's cores for a
  bounded sample of at most ~20 s; its objective is the FP64 reference for the
  reported objective gap.
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
    "config0": ("config0", "1D N=1024, 12 modes, gaps+5% dropout, lambda 0.1, tol 1e-8"),
    "config1": ("config1", "2D 255x255, 40 modes, r=3 discs on a 16-lattice, lambda 0.05, tol 1e-8"),
}


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


def cpu_baseline(problem, budget_s=20.0, max_runs=5):
    import numpy as np
    from oracle import ipm
    from oracle import operator as op
    workers = max(1, min(16, os.cpu_count() or 1))
    b = op.m_perpt_z(problem.z0, problem.dims, workers=workers)
    mt = lambda v: op.mtm(v, problem.dims, problem.missing, workers=workers)   # noqa: E731
    times = []
    res = None
    t_start = time.perf_counter()
    while len(times) < max_runs and (time.perf_counter() - t_start) < budget_s:
        t0 = time.perf_counter()
        res = ipm.solve(b, problem.lam, mt, ipm.IPMOptions(tol=1e-8))
        times.append(time.perf_counter() - t0)
    f_ref = ipm.objective(res.beta, np.abs(res.beta), mt(res.beta), b, problem.lam)
    med = statistics.median(times)
    return {"value": 1.0 / med, "unit": "solves/s", "cores": workers, "kind": "port",
            "sample": f"{len(times)} FP64 oracle solve(s) of the same instance "
                      f"(median {med:.3f} s, scipy pocketfft workers={workers}, "
                      f"{res.iterations} IPM iterations)"}, f_ref


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="config1", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every "
              "instance fits one MI355X; sharding is described in DESIGN.md)", file=sys.stderr)
        return 2

    import ctypes

    import numpy as np
    import torch

    import csipm.workloads as wl
    from csipm import _lib
    from csipm.fft_operator import FFTParameters, _ptr
    from csipm.model import FFTNLPModel, ipm_solve

    dev = torch.device("cuda:0")
    gen, desc = WORKLOADS[args.workload]
    problem = getattr(wl, gen)()
    params = FFTParameters(len(problem.dims), problem.dims, problem.z0, problem.lam,
                           problem.missing)
    nlp = FFTNLPModel(params, rdft=True, dtype=torch.float64)

    for _ in range(args.warmup):
        ipm_solve(nlp, tol=1e-8)
    torch.cuda.synchronize()
    lat = []
    stats = None
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        stats = ipm_solve(nlp, tol=1e-8)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
    torch.cuda.synchronize()
    total = time.perf_counter() - t_all

    # roofline of the dominant pass of the fused K apply
    # the K apply the CG runs (spectrum-ordered vectors of length U = 2*Nc)
    n = 2 * nlp.op_fft.Nc
    dt = nlp.op_fft.dtype
    rng = np.random.default_rng(0)
    x = torch.from_numpy(rng.standard_normal(2 * n)).to(dev, dt)
    L1 = torch.from_numpy(rng.uniform(0.1, 3, n)).to(dev, dt)
    L2 = torch.from_numpy(rng.uniform(-1, 1, n)).to(dev, dt)
    y = torch.empty(2 * n, dtype=dt, device=dev)
    ms = (ctypes.c_double * 8)()
    by = (ctypes.c_double * 8)()
    npass = ctypes.c_int()
    h = nlp.op_fft.handle
    _lib.call("csipm_kkt_apply_timed", h, _ptr(x), _ptr(y), _ptr(L1), _ptr(L2), 20, 1, ms, by,
              ctypes.byref(npass))
    _lib.call("csipm_kkt_apply_timed", h, _ptr(x), _ptr(y), _ptr(L1), _ptr(L2), 200, 1, ms, by,
              ctypes.byref(npass))
    passes = [(_lib.lib().csipm_kkt_pass_name(h, 1, i).decode(), ms[i], by[i])
              for i in range(npass.value)]
    name, pms, pbytes = max(passes, key=lambda p: p[1])
    achieved = pbytes / (pms * 1e-3) / 1e9
    kapply_ms = sum(p[1] for p in passes)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": load_traffic(args.workload, name),
                "kernel": name, "kernel_ms": round(pms, 5), "kernel_bytes": int(pbytes),
                "k_apply_ms": round(kapply_ms, 5),
                "passes": {p[0]: {"ms": round(p[1], 5), "bytes": int(p[2])} for p in passes}}

    out = {
        "metric": "solves/s at <=1e-6 obj gap, batched FP32-op vs plain FP64; p50/p99 latency",
        "value": round(args.steps / total, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{problem.name}: {desc}; plain FP64 one solve at a time",
                   "dims": list(problem.dims), "N": problem.N, "missing": int(len(problem.missing)),
                   "lambda": problem.lam},
        "p50_ms": round(1e3 * statistics.median(lat), 3),
        "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
        "ipm_iterations": stats.iterations,
        "cg_iterations": stats.cg_iterations,
        "objective": stats.objective,
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        cb, f_ref = cpu_baseline(problem)
        out["cpu_baseline"] = cb
        out["objective_ref"] = f_ref
        out["obj_rel_gap"] = abs(stats.objective - f_ref) / abs(f_ref)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
