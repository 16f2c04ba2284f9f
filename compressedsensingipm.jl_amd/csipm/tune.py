"""Autotuner of the batch engine's FFT pass launch shapes.

    python -m csipm.tune [--quick] [--part k/n] [--out PATH]   (from compressedsensingipm.jl_amd/)
    python -m csipm.tune --merge part1.json part2.json ... --out PATH

For every line-transform shape the serving and batch workloads meet (the odd
2D sides 63..511, the odd 3D sides 63..191 and the 255 / 255^3
configurations), it builds a batch of instances of that shape, and for each
pass kind (line pass, dimension-1 pass) times the batch's FP32 CG operator
apply (csipm_batch_apply_timed, HIP events) under the candidate launch
shapes: the tile capped at the built-in size and one or two halvings of it,
on 128 or 256 threads where the tile allows.  The fastest candidate that
beats the built-in choice by more than 3% is written to the table
(tune_table.json beside this module), keyed on (pass kind, element bytes,
FFT length, LDS rows), which `load()` hands to the library
(csipm_tune_set) whenever libcsipm is loaded.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tune_table.json")
KIND_NAMES = {0: "lines", 1: "dim1"}


def load(lib=None, path=TABLE):
    """Apply the committed table (no-op when it is absent).  Returns the
    number of entries applied."""
    if not os.path.exists(path):
        return 0
    if lib is None:
        from . import _lib
        lib = _lib.lib()
    with open(path) as f:
        tab = json.load(f)
    n = 0
    for e in tab.get("entries", []):
        rc = lib.csipm_tune_set(int(e["kind"]), int(e["elem_bytes"]), int(e["fft_len"]),
                                int(e["rows"]), int(e["lg_cap"]), int(e["nt"]))
        if rc == 0:
            n += 1
    return n


def fft_shape(op, axis):
    from . import _lib
    w, r, m = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _lib.call("csipm_op_fft_shape", op.handle, int(axis), ctypes.byref(w), ctypes.byref(r),
              ctypes.byref(m))
    return w.value, r.value, m.value


def _batch(dims, count):
    import csipm.workloads as wl
    from csipm.fft_operator import FFTOperator
    probs = [wl.batch_instance(dims, seed=1000 + k) for k in range(count)]
    ops = [FFTOperator(p.N, len(p.dims), p.dims, p.missing, mixed=True) for p in probs]
    return probs, ops


def _time(probs, ops, kind_tag, reps):
    """Summed HIP-event time (ms) of the batch's launches whose name holds
    kind_tag, fresh batch per call (tables rebuilt under the current table)."""
    from csipm.batch import BatchSolver, slots_of
    bs = BatchSolver(sum(slots_of(o) for o in ops), max_instances=len(ops), mixed=True)
    for p, o in zip(probs, ops):
        if bs.admit(o, p.z0, p.lam) is None:
            raise RuntimeError("tuning pool too small")
    bs.apply_timed(2)
    launches = bs.apply_timed(reps)
    bs.close()
    return sum(ms for name, ms, _ in launches if kind_tag(name))


def tune_shape(dims, count, reps, lib, log):
    """Entries for the passes of one shape (FP32)."""
    probs, ops = _batch(dims, count)
    d = len(dims)
    out = []
    kinds = [(1, 0)] + ([(0, 1)] if d >= 2 else [])
    for kind, axis in kinds:
        w, r, mode = fft_shape(ops[0], axis)
        if mode == 3:
            continue
        if kind == 1:
            tag = lambda n: ("dim1" in n and n.startswith(f"{d}d:"))          # noqa: E731
        else:
            tag = lambda n: ("lines" in n and n.startswith(f"{d}d:"))         # noqa: E731
        lib.csipm_tune_set(kind, 8, w, r, -1, 0)
        base = _time(probs, ops, tag, reps)
        best = (base, None)
        for cap in (3, 2, 1):
            for nt in (128, 256):
                lib.csipm_tune_set(kind, 8, w, r, cap, nt)
                t = _time(probs, ops, tag, reps)
                if t < best[0]:
                    best = (t, (cap, nt))
        lib.csipm_tune_set(kind, 8, w, r, -1, 0)
        gain = base / best[0] if best[1] else 1.0
        log(f"{'x'.join(map(str, dims))} {KIND_NAMES[kind]} W={w} R={r} mode={mode}: built-in "
            f"{base:.4f} ms, best {best[0]:.4f} ms {best[1]} ({gain:.3f}x)")
        if best[1] and gain > 1.03:
            out.append({"kind": kind, "elem_bytes": 8, "fft_len": w, "rows": r,
                        "lg_cap": best[1][0], "nt": best[1][1],
                        "shape": list(dims), "ms_builtin": round(base, 5),
                        "ms_tuned": round(best[0], 5)})
    del ops
    return out


def shapes(quick=False):
    s2 = list(range(63, 512, 2))
    s3 = list(range(63, 192, 2))
    if quick:
        s2 = s2[::16]
        s3 = s3[::8]
    return [(s, s) for s in s2] + [(s, s, s) for s in s3] + [(255, 255), (255, 255, 255)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true", help="every 16th 2D / 8th 3D side")
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--elems", type=float, default=2e7,
                    help="instances per shape: about this many samples per batch")
    ap.add_argument("--part", default="1/1", help="k/n: tune the k-th of n slices of the shapes")
    ap.add_argument("--dims", type=int, default=0,
                    help="tune only the shapes of this dimensionality (2 or 3)")
    ap.add_argument("--merge", nargs="*", default=None,
                    help="merge these partial tables into --out (no tuning)")
    args = ap.parse_args()
    if args.merge is not None:
        return merge(args.merge, args.out)
    from . import _lib
    import numpy as np
    lib = _lib.lib()
    lib.csipm_tune_clear()

    def log(m):
        print(m, flush=True)
    entries = []
    t0 = time.perf_counter()
    k, n = (int(v) for v in args.part.split("/"))
    todo = [d for d in dict.fromkeys(shapes(args.quick)) if not args.dims or len(d) == args.dims]
    for dims in todo[k - 1::n]:
        count = int(max(1, min(256, args.elems // int(np.prod(dims)))))
        entries += tune_shape(dims, count, args.reps, lib, log)
    # one entry per key (shapes sharing a key keep the larger gain)
    best = {}
    for e in entries:
        k = (e["kind"], e["elem_bytes"], e["fft_len"], e["rows"])
        g = e["ms_builtin"] / e["ms_tuned"]
        if k not in best or g > best[k]["ms_builtin"] / best[k]["ms_tuned"]:
            best[k] = e
    tab = {"generated_by": "python -m csipm.tune" + (" --quick" if args.quick else "") +
                           f" --part {args.part}" + (f" --dims {args.dims}" if args.dims else "") +
                           f" --elems {args.elems:g}",
           "device": _device_name(), "seconds": round(time.perf_counter() - t0, 1),
           "entries": sorted(best.values(), key=lambda e: (e["kind"], e["fft_len"], e["rows"]))}
    with open(args.out, "w") as f:
        json.dump(tab, f, indent=1)
    log(f"{len(tab['entries'])} entries -> {args.out}")
    return 0


def merge(paths, out):
    """One table from partial ones (the larger gain per key)."""
    best, meta = {}, []
    for p in paths:
        with open(p) as f:
            t = json.load(f)
        meta.append({k: t.get(k) for k in ("generated_by", "device", "seconds")})
        for e in t["entries"]:
            k = (e["kind"], e["elem_bytes"], e["fft_len"], e["rows"])
            g = e["ms_builtin"] / e["ms_tuned"]
            if k not in best or g > best[k]["ms_builtin"] / best[k]["ms_tuned"]:
                best[k] = e
    tab = {"generated_by": "python -m csipm.tune --merge", "parts": meta,
           "entries": sorted(best.values(), key=lambda e: (e["kind"], e["fft_len"], e["rows"]))}
    with open(out, "w") as f:
        json.dump(tab, f, indent=1)
    print(f"{len(tab['entries'])} entries -> {out}", flush=True)
    return 0


def _device_name():
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:
        return ""


if __name__ == "__main__":
    sys.exit(main())
