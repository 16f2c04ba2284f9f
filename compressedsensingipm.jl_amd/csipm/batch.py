"""Batched and continuously-batched solves over libcsipm's batch engine.

The reference solves one FFTNLPModel per MadNLP call
(src/CompressedSensingIPM.jl:31-36).  Here many problems of different
dimensionalities and sizes share every kernel launch of the interior-point
iteration: `BatchSolver` wraps the csipm_batch_* C ABI (admit / step / poll /
harvest), `solve_batch` runs a fixed set to completion, and `serve` runs an
arrival stream with continuous batching: requests are admitted between
interior-point iterations as pool slots (device memory) free up, while the
others keep iterating.
"""
from __future__ import annotations

import ctypes
import heapq
import itertools
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .fft_operator import FFTOperator, _ptr, _stream


def _prec(mixed):
    return _lib.MIXED if mixed else _lib.F64


def slots_of(op: FFTOperator) -> int:
    s = ctypes.c_int64()
    _lib.call("csipm_batch_slots", op.handle, ctypes.byref(s))
    return int(s.value)


def pool_bytes(slots: int, mixed: bool) -> int:
    b = ctypes.c_int64()
    _lib.call("csipm_batch_bytes", _prec(mixed), int(slots), ctypes.byref(b))
    return int(b.value)


def op_bytes(op: FFTOperator) -> int:
    """Device bytes the handle holds itself (csipm_op_bytes)."""
    b = ctypes.c_int64()
    _lib.call("csipm_op_bytes", op._h, ctypes.byref(b))
    return int(b.value)


def footprint(dims, n_missing, mixed=True):
    """(pool slots, handle device bytes) of a problem before its handle is
    made (csipm_op_footprint)."""
    dims = tuple(int(d) for d in dims)
    sizes = (ctypes.c_int64 * 3)(*(list(dims) + [1] * (3 - len(dims))))
    sl = ctypes.c_int64()
    by = ctypes.c_int64()
    _lib.call("csipm_op_footprint", len(dims), sizes, _prec(mixed), int(n_missing),
              ctypes.byref(sl), ctypes.byref(by))
    return int(sl.value), int(by.value)


@dataclass
class Result:
    beta: torch.Tensor
    objective: float
    iterations: int
    cg_iterations: int
    status: str
    dual_residual: float
    mu: float
    tag: object = None
    t_admit: float = 0.0
    t_done: float = 0.0
    error: str = ""


_STATUS = {0: "solved", 1: "max_iter", 2: "failure"}


def _rejected(tag, why, t):
    nan = float("nan")
    return Result(None, nan, 0, 0, "rejected", nan, nan, tag, t, t, why)


class BatchSolver:
    """One batch engine: a pool of `slots` spectrum-order slots for up to
    `max_instances` problems of precision F64 (mixed=False) or MIXED."""

    def __init__(self, slots, max_instances=256, mixed=False, device="cuda"):
        self.mixed = bool(mixed)
        self.device = torch.device(device)
        self.slots = int(slots)
        self.max_instances = int(max_instances)
        b = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.call("csipm_batch_create", ctypes.byref(b), _prec(mixed), self.slots,
                      self.max_instances, _stream())
        self._b = b
        self._live = {}      # id -> (op, z0, beta, tag, t_admit)

    def close(self):
        if getattr(self, "_b", None) is not None and self._b.value:
            _lib.lib().csipm_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, op: FFTOperator, z0, lam, tag=None, cg_itmax=0):
        """Admit a problem; returns its id, or None when the pool or the id
        table is full (the caller retries after harvesting).  A handle serves
        one live instance at a time: admitting one that is live raises."""
        if op.mixed != self.mixed:
            raise ValueError("operator precision differs from the batch")
        z = op._vec(z0, dtype=torch.float64)
        beta = torch.empty(op.N, dtype=torch.float64, device=op.device)
        i = ctypes.c_int()
        rc = _lib.lib().csipm_batch_admit(self._b, op.handle, _ptr(z), float(lam), _ptr(beta),
                                          int(cg_itmax), ctypes.byref(i))
        if rc == _lib.ERR_ALLOC:
            return None
        _lib.check(rc)
        self._live[i.value] = (op, z, beta, tag, time.perf_counter())
        return i.value

    def admit_many(self, items, cg_itmax=0):
        """Admit several problems in one call (csipm_batch_admit_many).
        items: (op, z, beta, lam, tag) with z and beta FP64 device vectors of
        op.N entries (z0 and the result buffer, e.g. from `buffers`).
        Returns one entry per item: its id, None when it does not fit now
        (later, smaller ones are still tried), or the CsipmError that
        admitting it raised (the others are unaffected)."""
        n = len(items)
        if n == 0:
            return []
        out = [None] * n
        ok = []
        for k, (op, *_) in enumerate(items):
            if op.mixed != self.mixed:
                out[k] = ValueError("operator precision differs from the batch")
            else:
                ok.append(k)
        m = len(ok)
        if m == 0:
            return out
        hs = (ctypes.c_void_p * m)(*[items[k][0].handle for k in ok])
        zs = (ctypes.c_void_p * m)(*[_ptr(items[k][1]) for k in ok])
        bv = (ctypes.c_void_p * m)(*[_ptr(items[k][2]) for k in ok])
        lams = (ctypes.c_double * m)(*[float(items[k][3]) for k in ok])
        ids = (ctypes.c_int * m)()
        rcs = (ctypes.c_int * m)()
        rc = _lib.lib().csipm_batch_admit_many(self._b, m, hs, zs, lams, bv, int(cg_itmax), ids,
                                               rcs)
        t = time.perf_counter()
        for q, k in enumerate(ok):
            op, z, beta, lam, tag = items[k]
            if ids[q] >= 0:
                self._live[ids[q]] = (op, z, beta, tag, t)
                out[k] = ids[q]
            elif rcs[q] not in (_lib.OK, _lib.ERR_ALLOC):
                out[k] = _lib.CsipmError(rcs[q], f"admission of problem {k} failed "
                                                 f"(code {rcs[q]})")
        _lib.check(rc)
        return out

    @staticmethod
    def buffers(op, z0):
        """(z, beta): z0 as an FP64 device vector and an empty result vector,
        the per-problem buffers admit_many takes."""
        return (op._vec(z0, dtype=torch.float64),
                torch.empty(op.N, dtype=torch.float64, device=op.device))

    def step(self, tol=1e-8, max_iter=200, cg_rtol=1e-10, cg_atol=0.0):
        """One interior-point iteration of every admitted, unfinished
        instance; returns the number still iterating.  cg_atol / cg_rtol: the
        FP64 PCG stopping rule (csipm.model.ipm_solve)."""
        r = ctypes.c_int()
        _lib.call("csipm_batch_step", self._b, float(tol), int(max_iter), float(cg_atol),
                  float(cg_rtol), ctypes.byref(r))
        return r.value

    def finished(self):
        cap = max(1, len(self._live))
        ids = (ctypes.c_int * cap)()
        n = ctypes.c_int()
        _lib.call("csipm_batch_poll", self._b, ids, cap, ctypes.byref(n))
        return [ids[k] for k in range(min(n.value, cap))]

    def harvest(self, i, retire=None) -> Result:
        """Results of finished instance i (its beta is written in stream
        order).  retire: a list that receives the instance's operator handle
        instead of letting it go here (a server destroys retired handles off
        its critical path)."""
        info = (ctypes.c_double * 8)()
        _lib.call("csipm_batch_harvest", self._b, int(i), info)
        op, z, beta, tag, t0 = self._live.pop(i)
        if retire is not None:
            retire.append(op)
        return Result(beta, info[0], int(info[1]), int(info[2]), _STATUS[int(info[3])], info[4],
                      info[5], tag, t0, time.perf_counter())

    def harvest_many(self, ids, retire=None):
        """harvest() of several finished instances in one call
        (csipm_batch_harvest_many)."""
        n = len(ids)
        if n == 0:
            return []
        idv = (ctypes.c_int * n)(*ids)
        info = (ctypes.c_double * (8 * n))()
        _lib.call("csipm_batch_harvest_many", self._b, n, idv, info)
        t = time.perf_counter()
        out = []
        for k, i in enumerate(ids):
            op, z, beta, tag, t0 = self._live.pop(i)
            if retire is not None:
                retire.append(op)
            f = info[8 * k:8 * k + 8]
            out.append(Result(beta, f[0], int(f[1]), int(f[2]), _STATUS[int(f[3])], f[4], f[5],
                              tag, t0, t))
        return out

    @property
    def live(self):
        return len(self._live)

    def stats(self):
        """Host-side accounting of the engine (seconds / counts)."""
        out = (ctypes.c_double * 8)()
        _lib.call("csipm_batch_stats", self._b, out)
        keys = ("step_s", "tables_s", "capture_s", "init_s", "cg_s", "table_builds",
                "captures", "steps")
        return {k: out[i] for i, k in enumerate(keys)}

    def apply_timed(self, nrep=20):
        """Per-launch HIP-event times of the batch's CG operator apply:
        list of (name, ms, algorithmic bytes)."""
        cap = 64
        ms = (ctypes.c_double * cap)()
        by = (ctypes.c_double * cap)()
        n = ctypes.c_int()
        _lib.call("csipm_batch_apply_timed", self._b, int(nrep), ms, by, cap, ctypes.byref(n))
        L = _lib.lib()
        return [(L.csipm_batch_launch_name(self._b, k).decode(), ms[k], by[k])
                for k in range(n.value)]


def solve_batch(ops, z0s, lams, mixed=False, tol=1e-8, max_iter=200, cg_rtol=1e-10,
                cg_atol=0.0):
    """Solve a fixed set of problems in one batch (all admitted up front).
    Returns the results in input order."""
    slots = sum(slots_of(op) for op in ops)
    bs = BatchSolver(slots, max_instances=max(1, len(ops)), mixed=mixed,
                     device=ops[0].device if ops else "cuda")
    ids = []
    for k, (op, z, lam) in enumerate(zip(ops, z0s, lams)):
        i = bs.admit(op, z, lam, tag=k)
        if i is None:
            raise RuntimeError("batch pool too small")
        ids.append(i)
    while bs.step(tol, max_iter, cg_rtol, cg_atol) > 0:
        pass
    out = [None] * len(ops)
    for i in bs.finished():
        r = bs.harvest(i)
        out[r.tag] = r
    bs.close()
    return out


@dataclass(order=True)
class Request:
    arrival: float
    seq: int
    problem: object = field(compare=False)


def serve(requests, make_op, mixed=True, budget_bytes=200 << 30, pool_fraction=0.75,
          max_instances=512, tol=1e-6, max_iter=200, cg_rtol=1e-10, cg_atol=0.0,
          clock=time.perf_counter, on_done=None, prep_workers=4, scan=64, stats=None,
          solver=None, lanes=None):
    """Continuous batching over an arrival stream.

    requests: iterable of Request (arrival time in seconds from the start,
    problem with .dims .z0 .lam .missing); make_op(problem) -> FFTOperator,
    called by `prep_workers` helper threads (0: inline).

    Device memory: one engine pool of pool_fraction * budget_bytes, made
    once, plus the handles of the requests being prepared or solved, which
    share the rest of the budget (csipm_op_footprint before a handle exists,
    csipm_op_bytes after).  A request whose slots exceed the pool, or whose
    handle exceeds the handle share, can never run: it comes back at once as
    a Result with status "rejected" (so does one whose handle cannot be made
    or admitted).  Handles are made in arrival order while their bytes fit;
    requests are admitted in arrival order as pool slots and ids free up,
    and up to `scan` later ready requests may pass one that does not fit.
    Every loop turn runs one interior-point iteration of all admitted
    problems and retires the finished ones (their handles are freed at
    once).  Returns the Results (tag = the request), admit / finish times
    relative to the start.  stats: a dict that receives the engine's
    accounting (BatchSolver.stats) and the loop's own split of the wall time
    (seconds in steps, admission, harvest, idle waits).

    solver: an empty BatchSolver of the same precision to serve with (a
    server's pool, made once at start-up and kept across calls; its slots
    replace pool_fraction * budget_bytes, the handles get the rest of the
    budget); by default serve makes one and frees it on return.

    lanes: instead of `solver`, several such engines, [(solver, accepts),
    ...] with accepts(problem) -> bool (None: any problem); a request goes
    to the first lane that accepts it.  Each turn the lanes' interior-point
    iterations run in parallel threads (the library releases the GIL), so
    one engine's host work (synchronisations, launch-table builds) overlaps
    the other's kernels, and the lanes' vector kernels run on their own
    streams: a lane of 3D problems and one of 1D/2D problems keep the small
    problems' vector work off the 3D problems' stream."""
    from collections import deque
    from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor
    from concurrent.futures import wait as fwait
    per_slot = pool_bytes(1, mixed)
    if lanes is not None and solver is not None:
        raise ValueError("serve takes a solver or lanes, not both")
    if solver is not None:
        lanes = [(solver, None)]
    if lanes is not None:
        lanes = [(s, acc) for s, acc in lanes]
        if not lanes:
            raise ValueError("serve needs at least one lane")
        for s, _ in lanes:
            if s.mixed != bool(mixed) or s.live:
                raise ValueError("serve needs empty solvers of the requested precision")
    solvers = [s for s, _ in lanes] if lanes is not None else []
    pool_slots = sum(s.slots for s in solvers) if solvers else \
        max(1, int(budget_bytes * pool_fraction) // per_slot)
    handle_budget = budget_bytes - pool_slots * per_slot
    pending = list(requests)
    heapq.heapify(pending)
    waiting = deque()   # arrived, not admitted: [req, future or op, bytes held, lane]
    nsub = 0            # waiting requests (a prefix) whose handles are made or in the making
    out = []
    held = 0            # handle bytes of waiting (prepared) and live requests
    live_bytes = {}     # request seq -> handle bytes
    dev = torch.cuda.current_device()

    def prepare(problem):
        # the handle and the request's device buffers (z0 in FP64, beta),
        # made off the loop's critical path
        with torch.cuda.device(dev):
            op = make_op(problem)
            try:
                z, beta = BatchSolver.buffers(op, problem.z0)
                nb = op_bytes(op)
            except Exception:
                op.close()
                raise
            return op, z, beta, nb

    def finish(r):
        out.append(r)
        if on_done:
            on_done(r)

    prep = ThreadPoolExecutor(max_workers=prep_workers) if prep_workers > 0 else None
    own = not solvers
    if own:
        solvers = [BatchSolver(pool_slots, max_instances=max_instances, mixed=mixed)]
        lanes = [(solvers[0], None)]
    stepper = ThreadPoolExecutor(max_workers=len(solvers)) if len(solvers) > 1 else None

    def lane_of(problem):
        for k, (_, acc) in enumerate(lanes):
            if acc is None or acc(problem):
                return k
        return -1

    t_start = clock()
    acc = {"step_s": 0.0, "admit_s": 0.0, "harvest_s": 0.0, "idle_s": 0.0, "steps": 0,
           "live_sum": 0}
    closing = []        # helper-thread destructions of retired handles, not yet done
    drained = False
    try:
        while pending or waiting or any(s.live for s in solvers):
            now = clock() - t_start
            while pending and pending[0].arrival <= now:
                req = heapq.heappop(pending)
                p = req.problem
                try:
                    slots, hb = footprint(p.dims, len(p.missing), mixed)
                except _lib.CsipmError as e:
                    finish(_rejected(req, str(e), now))
                    continue
                k = lane_of(p)
                if k < 0:
                    finish(_rejected(req, "no lane takes this problem", now))
                    continue
                if slots > solvers[k].slots or hb > handle_budget:
                    finish(_rejected(req, f"needs {slots} pool slots (pool {solvers[k].slots}) "
                                          f"and {hb} handle bytes (share {handle_budget})", now))
                    continue
                waiting.append([req, None, hb, k])
            # make handles, in arrival order, while their bytes fit (the
            # first nsub waiting requests have theirs made or in the making)
            for w in itertools.islice(waiting, nsub, None):
                if held + w[2] > handle_budget:
                    break
                held += w[2]
                nsub += 1
                w[1] = prep.submit(prepare, w[0].problem) if prep is not None else \
                    _made(prepare, w[0].problem)
            # admit ready requests in arrival order, in one call
            t_a = time.perf_counter()
            skipped = 0
            scanned = 0
            cand = []          # (waiting entry, op, z, beta, bytes)
            # more than this could not be admitted to a lane
            room = [s.max_instances - s.live + scan for s in solvers]
            ncand = [0] * len(solvers)
            full = 0           # passed over because their lane has no room
            for w in waiting:
                if w[1] is None or skipped > scan or full > 8 * scan:
                    break
                if ncand[w[3]] >= room[w[3]]:
                    if len(solvers) == 1:
                        break
                    scanned += 1
                    full += 1
                    continue
                scanned += 1
                if not w[1].done():
                    skipped += 1
                    continue
                ncand[w[3]] += 1
                try:
                    op, z, beta, nb = w[1].result()
                except (_lib.CsipmError, ValueError, RuntimeError) as e:
                    w[1] = _Ready(None, e)
                    cand.append((w, None, None, None, 0))
                    continue
                cand.append((w, op, z, beta, nb))
            todo = [[(op, z, beta, w[0].problem.lam, w[0])
                     for w, op, z, beta, _ in cand if op is not None and w[3] == k]
                    for k in range(len(solvers))]
            keys = [[id(w) for w, op, *_ in cand if op is not None and w[3] == k]
                    for k in range(len(solvers))]
            res = {}
            outs = {}
            t_s = None
            if len(solvers) == 1:
                if todo[0]:
                    res.update(zip(keys[0], solvers[0].admit_many(todo[0])))
                lane_live = [solvers[0].live]
            else:
                # each lane's admissions, interior-point iteration and harvest
                # in its own thread: one lane's host work overlaps the other's
                # kernels
                run = [k for k in range(len(solvers)) if todo[k] or solvers[k].live]
                t_s = time.perf_counter()
                futs = [(k, stepper.submit(_lane_turn, solvers[k], todo[k], tol, max_iter,
                                           cg_rtol, cg_atol)) for k in run]
                outs = {k: f.result() for k, f in futs}
                for k, (got, live, fin, ret) in outs.items():
                    res.update(zip(keys[k], got))
                lane_live = [outs[k][1] if k in outs else solvers[k].live
                             for k in range(len(solvers))]
            gone = set()
            for w, op, z, beta, nb in cand:
                req = w[0]
                if op is not None:
                    held += nb - w[2]
                    w[2] = nb
                g = res.get(id(w)) if op is not None else w[1].exc
                if isinstance(g, Exception):
                    held -= w[2]
                    gone.add(id(w))
                    if op is not None:
                        op.close()
                    finish(_rejected(req, str(g), clock() - t_start))
                elif g is None:
                    if lane_live[w[3]] == 0:
                        # an empty engine cannot take it: it never will
                        held -= w[2]
                        gone.add(id(w))
                        op.close()
                        finish(_rejected(req, "does not fit an empty engine",
                                         clock() - t_start))
                    else:
                        w[1] = _Ready((op, z, beta, nb))
                else:
                    live_bytes[req.seq] = w[2]
                    gone.add(id(w))
            if gone:
                # the removed requests are among the first `scanned`
                head = [waiting.popleft() for _ in range(scanned)]
                for w in reversed(head):
                    if id(w) not in gone:
                        waiting.appendleft(w)
                nsub -= len(gone)
            stepped = [k for k in outs if outs[k][1]]
            if t_s is None:
                acc["admit_s"] += time.perf_counter() - t_a
            else:
                acc["admit_s"] += t_s - t_a
            if not stepped and not any(s.live for s in solvers):
                t_i = time.perf_counter()
                futs = [w[1] for w in waiting if w[1] is not None and not w[1].done()]
                nxt = pending[0].arrival - (clock() - t_start) if pending else None
                if futs:
                    fwait(futs, timeout=nxt if nxt is not None and nxt > 0 else None,
                          return_when=FIRST_COMPLETED)
                elif nxt is not None and nxt > 0:
                    time.sleep(nxt)
                acc["idle_s"] += time.perf_counter() - t_i
                continue
            retired = []
            if len(solvers) == 1:
                t_s = time.perf_counter()
                acc["live_sum"] += solvers[0].live
                solvers[0].step(tol, max_iter, cg_rtol, cg_atol)
                t_h = time.perf_counter()
                fin = solvers[0].harvest_many(solvers[0].finished(), retire=retired)
            else:
                # (lanes: the turn's wall time counts as step time, harvests included)
                t_h = time.perf_counter()
                acc["live_sum"] += sum(outs[k][1] for k in stepped)
                fin = [r for k in stepped for r in outs[k][2]]
                for k in stepped:
                    retired += outs[k][3]
            acc["step_s"] += t_h - t_s
            acc["steps"] += 1
            for r in fin:
                r.t_admit -= t_start
                r.t_done = clock() - t_start
                held -= live_bytes.pop(r.tag.seq, 0)
                finish(r)
            # the retired handles are destroyed on a helper thread, beside the
            # next step (destruction synchronises the handle's stream)
            if retired:
                if prep is not None:
                    closing = [f for f in closing if not f.done()]
                    closing.append(prep.submit(_close_all, retired))
                else:
                    _close_all(retired)
            acc["harvest_s"] += time.perf_counter() - t_h
        drained = True
    finally:
        if stats is not None:
            try:
                tot = {}
                for s in solvers:
                    for k, v in s.stats().items():
                        tot[k] = tot.get(k, 0.0) + v
                stats.update(tot)
            except Exception:
                pass
            stats.update({"loop_" + k: v for k, v in acc.items()})
            stats["wall_s"] = clock() - t_start
        t_c = time.perf_counter()
        if own:
            solvers[0].close()
        if stepper is not None:
            stepper.shutdown(wait=True)
        if prep is not None:
            if drained:
                # the last turns' retired handles are destroyed before return
                # (pending handle preparations are cancelled only on an error)
                fwait(closing)
            prep.shutdown(wait=True, cancel_futures=True)
        if stats is not None:
            stats["shutdown_s"] = time.perf_counter() - t_c
    return out


def _lane_turn(s, items, tol, max_iter, cg_rtol, cg_atol):
    """One serving turn of a lane (run in a lane thread): its admissions, one
    interior-point iteration of its instances and the finished ones'
    results.  Returns (admission results, live instances after admission,
    results, retired handles)."""
    got = s.admit_many(items) if items else []
    live = s.live
    fin, retired = [], []
    if live:
        s.step(tol, max_iter, cg_rtol, cg_atol)
        fin = s.harvest_many(s.finished(), retire=retired)
    return got, live, fin, retired


class _Ready:
    """A made handle (or the error of making it) held in the waiting queue;
    quacks like a done future."""

    def __init__(self, op, exc=None):
        self.op = op
        self.exc = exc

    def done(self):
        return True

    def result(self):
        if self.exc is not None:
            raise self.exc
        return self.op


def _close_all(ops):
    for op in ops:
        op.close()


def _made(fn, *args):
    try:
        return _Ready(fn(*args))
    except (_lib.CsipmError, ValueError, RuntimeError) as e:
        return _Ready(None, e)
