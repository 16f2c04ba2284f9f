"""Operator microbenchmark: the fused condensed-KKT apply at a given shape.

    python compressedsensingipm.jl_amd/csipm/perf.py --dims 255 255 255 --dtype f32
    python compressedsensingipm.jl_amd/csipm/perf.py --copy-mib 1024   (copy calibration)
    python compressedsensingipm.jl_amd/csipm/perf.py --batch 64 --workload 2d255 --mixed

Prints one JSON line with the per-pass HIP-event times (csipm_kkt_apply_timed),
the algorithmic bytes of each pass and the achieved fraction of 8 TB/s.  The
copy mode runs one plain device-to-device copy of known size, the reference
point used to scale rocprofv3's FETCH_SIZE / WRITE_SIZE counters.
"""
import argparse
import ctypes
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dims", type=int, nargs="+", default=[255, 255, 255])
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f32")
    ap.add_argument("--reps", type=int, default=50)
    ap.add_argument("--missing-frac", type=float, default=0.05)
    ap.add_argument("--copy-mib", type=int, default=0)
    ap.add_argument("--layout", choices=["spectral", "beta", "reduced"], default="spectral",
                    help="spectral: the condensed-K apply inside CG (spectrum-ordered "
                         "vectors); beta: csipm_kkt_apply on reference-ordered vectors; "
                         "reduced: the (M^T M + D) apply of the mixed-precision CG")
    ap.add_argument("--mixed", action="store_true", help="FP32 operator of a mixed handle")
    ap.add_argument("--batch", type=int, default=0,
                    help="time the batch engine's CG apply over this many instances instead")
    ap.add_argument("--workload", default="", help="with --batch: 2d255 instances "
                    "(else batch_instance problems of --dims)")
    args = ap.parse_args()

    import numpy as np
    import torch

    dev = torch.device("cuda:0")
    if args.copy_mib:
        n = args.copy_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        # the achievable-bandwidth reference: timed copies of the same size
        # (bytes read + bytes written per copy)
        reps = 10
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        b.copy_(a)
        e0.record()
        for _ in range(reps):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        print(json.dumps({"copy_bytes": 4 * n, "copy_ms": round(ms, 4),
                          "copy_GBps_read_plus_write": round(2 * 4 * n / (ms * 1e-3) / 1e9, 1)}),
              flush=True)
        return 0

    from csipm import _lib
    from csipm.fft_operator import FFTOperator, _ptr

    if args.batch:
        return batch_main(args)
    dims = tuple(args.dims)
    N = int(np.prod(dims))
    rng = np.random.default_rng(0)
    miss = np.sort(rng.choice(N, size=int(args.missing_frac * N), replace=False))
    dt = torch.float64 if args.dtype == "f64" else torch.float32
    op = FFTOperator(N, len(dims), dims, miss, dtype=dt, mixed=args.mixed)
    dt = op.dtype
    layout = {"beta": 0, "spectral": 1, "reduced": 2}[args.layout]
    n = N if layout == 0 else 2 * op.Nc
    x = torch.randn(2 * n, dtype=dt, device=dev)
    L1 = torch.rand(n, dtype=dt, device=dev) + 0.1
    L2 = torch.rand(n, dtype=dt, device=dev) - 0.5
    y = torch.empty(2 * n, dtype=dt, device=dev)
    ms = (ctypes.c_double * 8)()
    by = (ctypes.c_double * 8)()
    npass = ctypes.c_int()
    h = op.handle
    _lib.call("csipm_kkt_apply_timed", h, _ptr(x), _ptr(y), _ptr(L1), _ptr(L2), 3, layout, ms,
              by, ctypes.byref(npass))
    _lib.call("csipm_kkt_apply_timed", h, _ptr(x), _ptr(y), _ptr(L1), _ptr(L2), args.reps,
              layout, ms, by, ctypes.byref(npass))
    passes = {}
    tot_ms = 0.0
    tot_b = 0.0
    for i in range(npass.value):
        name = _lib.lib().csipm_kkt_pass_name(h, layout, i).decode()
        gbs = by[i] / (ms[i] * 1e-3) / 1e9
        passes[name] = {"ms": round(ms[i], 5), "bytes": int(by[i]), "GB/s": round(gbs, 1),
                        "frac": round(gbs / HBM_PEAK_GBS, 4)}
        tot_ms += ms[i]
        tot_b += by[i]
    s = 8 if dt == torch.float64 else 4
    out = {"dims": list(dims), "dtype": "f64" if s == 8 else "f32", "layout": args.layout,
           "N": N, "U": 2 * op.Nc, "reps": args.reps,
           "k_apply_ms": round(tot_ms, 5), "bytes": int(tot_b),
           "k_apply_GB/s": round(tot_b / (tot_ms * 1e-3) / 1e9, 1),
           "k_apply_frac": round(tot_b / (tot_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "passes": passes}
    print(json.dumps(out), flush=True)
    return 0


def batch_main(args):
    """Per-launch times of the batch engine's CG operator apply."""
    import csipm.workloads as wl
    from csipm.batch import BatchSolver, slots_of
    from csipm.fft_operator import FFTOperator
    if args.workload == "2d255":
        probs = [wl.punched_2d_255(instance=i + 1) for i in range(args.batch)]
    else:
        probs = [wl.batch_instance(tuple(args.dims), seed=i + 1) for i in range(args.batch)]
    mixed = args.mixed or args.dtype == "f32"
    ops = [FFTOperator(p.N, len(p.dims), p.dims, p.missing, mixed=mixed) for p in probs]
    bs = BatchSolver(sum(slots_of(o) for o in ops), max_instances=len(ops), mixed=mixed)
    for p, o in zip(probs, ops):
        bs.admit(o, p.z0, p.lam)
    bs.apply_timed(3)
    launches = bs.apply_timed(args.reps)
    tot_ms = sum(t[1] for t in launches)
    tot_b = sum(t[2] for t in launches)
    out = {"batch": args.batch, "workload": args.workload or f"batch_instance{tuple(args.dims)}",
           "precision": "mixed" if mixed else "f64", "reps": args.reps,
           "apply_ms": round(tot_ms, 5), "bytes": int(tot_b),
           "apply_GB/s": round(tot_b / (tot_ms * 1e-3) / 1e9, 1),
           "launches": {n: {"ms": round(ms, 5), "bytes": int(b),
                            "GB/s": round(b / (ms * 1e-3) / 1e9, 1),
                            "frac": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
                        for n, ms, b in launches}}
    bs.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
