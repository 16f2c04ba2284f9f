"""Autotuner of the batch engine's FFT pass launch shapes.

    python -m csipm.tune [--quick] [--part k/n] [--out PATH]   (from compressedsensingipm.jl_amd/)
    python -m csipm.tune --merge part1.json part2.json ... --out PATH
    python -m csipm.tune --bluestein [--dims 3] [--out PATH]

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

--bluestein chooses the convolution length M of every Bluestein line length
of those shapes: the smallest factorable M >= 2L - 1 and the cheapest few
in [2L - 1, 2.25 L] by a stage-cost model are each timed (the shape's FP32
apply at its best launch shapes), and an M that beats the smallest by more
than 3% is added to the table's "bluestein" list (csipm_tune_bluestein),
with the launch-shape entries of its passes; the rest of the table is kept.
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
    for e in tab.get("bluestein", []):
        if lib.csipm_tune_bluestein(int(e["line_len"]), int(e["conv_len"])) == 0:
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


def radix_stages(m):
    """Radices of the library's plan of length m (its factorisation order:
    31, 29, 23, 19, then 16, 8, 4, 2, then 17, 13, 11, 7, 15, 9, 5, 3), or
    None when m does not factor."""
    out = []
    for r in (31, 29, 23, 19, 16, 8, 4, 2, 17, 13, 11, 7, 15, 9, 5, 3):
        while m % r == 0:
            out.append(r)
            m //= r
    return out if m == 1 else None


# relative cost of one radix-r stage per element (least-squares fit to timed
# batched applies of 3D Bluestein cubes on MI355X: a stage costs about the
# same whatever its radix, 8 and 16 less, the large odd radices more)
STAGE_COST = {2: 4.6, 3: 4.6, 4: 4.4, 5: 4.2, 7: 4.3, 8: 3.1, 9: 4.1, 11: 4.3, 13: 5.4,
              15: 6.4, 16: 3.6, 17: 4.3, 19: 5.1, 23: 5.8, 29: 7.8, 31: 6.1}


def bluestein_candidates(L, keep=5):
    """(smallest factorable M >= 2L - 1, [up to `keep` cheapest M in
    [2L - 1, 2.25 L] by the stage-cost model])."""
    ms = [m for m in range(2 * L - 1, int(2.25 * L) + 1) if radix_stages(m)]
    if not ms:
        return None, []
    cost = lambda m: m * sum(STAGE_COST[r] for r in radix_stages(m))   # noqa: E731
    return ms[0], sorted(ms, key=cost)[:keep]


def _shape_best(dims, count, reps, lib):
    """Best FP32 apply time (ms) of each pass kind of the shape under the
    launch-shape candidates: ({kind: (ms, (cap, nt) or None, built-in ms)},
    shapes)."""
    probs, ops = _batch(dims, count)
    d = len(dims)
    kinds = [(1, 0)] + ([(0, 1)] if d >= 2 else [])
    shp = {kind: fft_shape(ops[0], axis) for kind, axis in kinds}
    best = {}
    for cand in [None, (3, 128), (3, 256), (2, 128), (2, 256), (1, 128), (1, 256)]:
        for kind, _ in kinds:
            w, r, _m = shp[kind]
            lib.csipm_tune_set(kind, 8, w, r, *(cand or (-1, 0)))
        from csipm.batch import BatchSolver, slots_of
        bs = BatchSolver(sum(slots_of(o) for o in ops), max_instances=len(ops), mixed=True)
        for p, o in zip(probs, ops):
            if bs.admit(o, p.z0, p.lam) is None:
                raise RuntimeError("tuning pool too small")
        bs.apply_timed(2)
        launches = bs.apply_timed(reps)
        bs.close()
        for kind, _ in kinds:
            tag = "dim1" if kind == 1 else "lines"
            t = sum(ms for name, ms, _ in launches if tag in name and name.startswith(f"{d}d:"))
            if cand is None:
                base = t
                best[kind] = (t, cand, base)
            elif t < best[kind][0]:
                best[kind] = (t, cand, best[kind][2])
    for kind, _ in kinds:
        w, r, _m = shp[kind]
        lib.csipm_tune_set(kind, 8, w, r, -1, 0)
    for o in ops:
        o.close()
    return best, shp


def tune_bluestein(dims, count, reps, lib, log):
    """(bluestein entry or None, launch-shape entries) of a shape whose
    sides all equal L and take Bluestein's form."""
    L = dims[0]
    m0, cands = bluestein_candidates(L)
    if m0 is None:
        return None, []
    res = {}
    for m in dict.fromkeys([m0] + cands):
        lib.csipm_tune_bluestein(L, m)
        best, shp = _shape_best(dims, count, reps, lib)
        res[m] = (sum(v[0] for v in best.values()), best, shp)
    lib.csipm_tune_bluestein(L, 0)
    mb = min(res, key=lambda m: res[m][0])
    gain = res[m0][0] / res[mb][0]
    log(f"{'x'.join(map(str, dims))} Bluestein L={L}: " +
        ", ".join(f"M={m} {radix_stages(m)} {res[m][0]:.4f} ms" for m in res) +
        f" -> M={mb} ({gain:.3f}x)")
    if mb == m0 or gain <= 1.03:
        return None, []
    total, best, shp = res[mb]
    out = []
    for kind, (t, cand, base) in best.items():
        if cand is None or base / t <= 1.03:
            continue
        w, r, _m = shp[kind]
        out.append({"kind": kind, "elem_bytes": 8, "fft_len": w, "rows": r,
                    "lg_cap": cand[0], "nt": cand[1], "shape": list(dims),
                    "ms_builtin": round(base, 5), "ms_tuned": round(t, 5)})
    return ({"line_len": L, "conv_len": mb, "shape": list(dims),
             "ms_smallest": round(res[m0][0], 5), "ms_tuned": round(total, 5)}, out)


def main_bluestein(args, lib, log):
    """Bluestein lengths of the shapes' sides, merged into the table at --out
    (its other entries kept; entries of the same keys replaced)."""
    import numpy as np
    t0 = time.perf_counter()
    blue, launch = [], []
    for dims in dict.fromkeys(shapes(args.quick)):
        if args.dims and len(dims) != args.dims:
            continue
        if len(set(dims)) != 1 or radix_stages(dims[0]) is not None:
            continue
        L = dims[0]
        if L > 2 and all(L % q for q in range(2, int(L ** 0.5) + 1)) and radix_stages(L - 1):
            continue    # Rader's form
        count = int(max(1, min(256, args.elems // int(np.prod(dims)))))
        b, entries = tune_bluestein(dims, count, args.reps, lib, log)
        if b:
            blue.append(b)
            launch += entries
    tab = {"entries": [], "bluestein": []}
    if os.path.exists(args.out):
        with open(args.out) as f:
            tab = json.load(f)
    keys = {(e["kind"], e["elem_bytes"], e["fft_len"], e["rows"]) for e in launch}
    tab["entries"] = sorted([e for e in tab.get("entries", [])
                             if (e["kind"], e["elem_bytes"], e["fft_len"], e["rows"]) not in keys]
                            + launch, key=lambda e: (e["kind"], e["fft_len"], e["rows"]))
    lens = {b["line_len"] for b in blue}
    tab["bluestein"] = sorted([b for b in tab.get("bluestein", []) if b["line_len"] not in lens]
                              + blue, key=lambda b: b["line_len"])
    tab["generated_by_bluestein"] = ("python -m csipm.tune --bluestein" +
                                     (f" --dims {args.dims}" if args.dims else "") +
                                     f" --elems {args.elems:g}")
    tab["seconds_bluestein"] = round(time.perf_counter() - t0, 1)
    with open(args.out, "w") as f:
        json.dump(tab, f, indent=1)
    log(f"{len(blue)} Bluestein lengths, {len(launch)} launch entries -> {args.out}")
    return 0


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
    ap.add_argument("--bluestein", action="store_true",
                    help="choose the Bluestein convolution lengths (merged into --out)")
    args = ap.parse_args()
    if args.merge is not None:
        return merge(args.merge, args.out)
    from . import _lib
    import numpy as np
    lib = _lib.lib()
    lib.csipm_tune_clear()

    def log(m):
        print(m, flush=True)
    if args.bluestein:
        return main_bluestein(args, lib, log)
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
    blue = {}
    for p in paths:
        with open(p) as f:
            for b in json.load(f).get("bluestein", []):
                blue[b["line_len"]] = b
    tab = {"generated_by": "python -m csipm.tune --merge", "parts": meta,
           "entries": sorted(best.values(), key=lambda e: (e["kind"], e["fft_len"], e["rows"])),
           "bluestein": [blue[k] for k in sorted(blue)]}
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
