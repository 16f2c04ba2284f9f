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


_STATUS = {0: "solved", 1: "max_iter", 2: "failure"}


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
        table is full (the caller retries after harvesting)."""
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

    def step(self, tol=1e-8, max_iter=200, cg_rtol=1e-10):
        r = ctypes.c_int()
        _lib.call("csipm_batch_step", self._b, float(tol), int(max_iter), float(cg_rtol),
                  ctypes.byref(r))
        return r.value

    def finished(self):
        cap = max(1, len(self._live))
        ids = (ctypes.c_int * cap)()
        n = ctypes.c_int()
        _lib.call("csipm_batch_poll", self._b, ids, cap, ctypes.byref(n))
        return [ids[k] for k in range(min(n.value, cap))]

    def harvest(self, i) -> Result:
        info = (ctypes.c_double * 8)()
        _lib.call("csipm_batch_harvest", self._b, int(i), info)
        op, z, beta, tag, t0 = self._live.pop(i)
        return Result(beta, info[0], int(info[1]), int(info[2]), _STATUS[int(info[3])], info[4],
                      info[5], tag, t0, time.perf_counter())

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


def solve_batch(ops, z0s, lams, mixed=False, tol=1e-8, max_iter=200, cg_rtol=1e-10):
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
    while bs.step(tol, max_iter, cg_rtol) > 0:
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


def serve(requests, make_op, mixed=True, pool_budget_bytes=200 << 30, max_instances=512,
          tol=1e-6, max_iter=200, cg_rtol=1e-10, clock=time.perf_counter, on_done=None,
          prep_workers=4):
    """Continuous batching over an arrival stream.

    requests: iterable of Request (arrival time in seconds from the start,
    problem with .dims .z0 .lam .missing); make_op(problem) -> FFTOperator,
    called by `prep_workers` helper threads as requests arrive (0: inline).
    A request is admitted once it has arrived, its handle is ready and its
    slots fit the pool (sized from pool_budget_bytes); every loop turn runs
    one interior-point iteration of all admitted problems and retires the
    finished ones.  Returns the list of Results (tag = the request) with
    admit / finish times relative to the start."""
    from concurrent.futures import ThreadPoolExecutor
    slots_budget = max(1, pool_budget_bytes // max(1, pool_bytes(1, mixed)))
    pending = list(requests)
    heapq.heapify(pending)
    bs = None
    waiting = []
    out = []
    t_start = clock()
    ops, futs = {}, {}
    dev = torch.cuda.current_device()

    def prepare(problem):
        with torch.cuda.device(dev):
            return make_op(problem)

    prep = ThreadPoolExecutor(max_workers=prep_workers) if prep_workers > 0 else None
    while pending or waiting or (bs is not None and bs.live):
        now = clock() - t_start
        while pending and pending[0].arrival <= now:
            req = heapq.heappop(pending)
            waiting.append(req)
            if prep is not None:
                futs[req.seq] = prep.submit(prepare, req.problem)
        if bs is None:
            bs = BatchSolver(slots_budget, max_instances=max_instances, mixed=mixed)
        # admit in arrival order while the pool has room
        still = []
        for req in waiting:
            if req.seq not in ops:
                f = futs.get(req.seq)
                if f is not None and not f.done():
                    still.append(req)
                    continue
                ops[req.seq] = futs.pop(req.seq).result() if f is not None else \
                    make_op(req.problem)
            op = ops[req.seq]
            i = bs.admit(op, req.problem.z0, req.problem.lam, tag=req)
            if i is None:
                still.append(req)
        waiting = still
        if bs.live == 0:
            if pending or waiting:
                nxt = pending[0].arrival - (clock() - t_start) if pending else 0.0
                time.sleep(max(0.0002, min(0.001, nxt)))
            continue
        bs.step(tol, max_iter, cg_rtol)
        for i in bs.finished():
            r = bs.harvest(i)
            r.t_admit -= t_start
            r.t_done = clock() - t_start
            ops.pop(r.tag.seq, None)
            out.append(r)
            if on_done:
                on_done(r)
    if bs is not None:
        bs.close()
    if prep is not None:
        prep.shutdown()
    return out
