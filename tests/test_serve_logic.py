"""Host logic of csipm.batch.serve (continuous batching, lanes) on the CPU:
the engines are stand-ins that keep each admitted problem for a fixed number
of interior-point steps, so admission order, pool and id limits, the handle
budget, lane routing, rejections and handle retirement are checked without
a GPU.  Every request must come back exactly once and every handle made must
be closed exactly once."""
import contextlib
import itertools
import threading
import time

import pytest
import torch

import csipm.batch as B


class FakeOp:
    def __init__(self, problem):
        self.problem = problem
        self.slots = problem.slots
        self.steps = problem.steps
        self.nbytes = problem.nbytes
        self.closed = 0

    def close(self):
        self.closed += 1


class Prob:
    _seq = itertools.count()

    def __init__(self, dims, slots, steps, nbytes=1 << 20):
        self.dims = dims
        self.slots = slots
        self.steps = steps
        self.nbytes = nbytes
        self.missing = ()
        self.lam = 0.1
        self.z0 = None
        self.uid = next(Prob._seq)


class FakeSolver:
    """The BatchSolver surface serve() uses: a pool of `slots`, at most
    `max_instances` live problems, each finished after its step count."""

    def __init__(self, slots, max_instances):
        self.slots = slots
        self.max_instances = max_instances
        self.mixed = True
        self.used = 0
        self.ids = itertools.count()
        self._live = {}
        self.served = []
        self.threads = set()
        self.peak = 0

    @property
    def live(self):
        return len(self._live)

    def admit_many(self, items):
        out = []
        for op, z, beta, lam, tag in items:
            if self.live >= self.max_instances or self.used + op.slots > self.slots:
                out.append(None)
                continue
            i = next(self.ids)
            self._live[i] = [op.steps, op, tag, time.perf_counter()]
            self.used += op.slots
            out.append(i)
        self.peak = max(self.peak, self.live)
        return out

    def step(self, tol, max_iter, cg_rtol, cg_atol):
        self.threads.add(threading.get_ident())
        for v in self._live.values():
            v[0] -= 1
        return sum(v[0] > 0 for v in self._live.values())

    def finished(self):
        return [i for i, v in self._live.items() if v[0] <= 0]

    def harvest_many(self, ids, retire=None):
        res = []
        for i in ids:
            _, op, tag, t0 = self._live.pop(i)
            self.used -= op.slots
            self.served.append(tag.problem.uid)
            if retire is not None:
                retire.append(op)
            res.append(B.Result(None, 0.0, op.steps, 0, "solved", 0.0, 0.0, tag, t0,
                                time.perf_counter()))
        return res

    def stats(self):
        return {"steps": 0.0}


@pytest.fixture
def fake_env(monkeypatch):
    made = []

    def make_op(problem):
        op = FakeOp(problem)
        made.append(op)
        return op
    monkeypatch.setattr(B, "footprint", lambda dims, nmiss, mixed=True: (
        _FOOT[dims][0], _FOOT[dims][1]))
    monkeypatch.setattr(B, "op_bytes", lambda op: op.nbytes)
    monkeypatch.setattr(B, "pool_bytes", lambda slots, mixed: 100 * slots)
    monkeypatch.setattr(B.BatchSolver, "buffers", staticmethod(lambda op, z0: (None, None)))
    monkeypatch.setattr(torch.cuda, "current_device", lambda: 0)
    monkeypatch.setattr(torch.cuda, "device", lambda d: contextlib.nullcontext())
    return make_op, made


_FOOT = {}


def _stream(n, seed=0):
    """n problems of three classes (dims tell the class), varied sizes and
    step counts."""
    import random
    rnd = random.Random(seed)
    probs = []
    for k in range(n):
        u = rnd.random()
        if u < 0.7:
            dims = (1000 + rnd.randrange(64),)
            slots, steps = 1000 + rnd.randrange(1000), rnd.randrange(3, 9)
        elif u < 0.95:
            dims = (50 + rnd.randrange(8), 50)
            slots, steps = 5000 + rnd.randrange(5000), rnd.randrange(4, 10)
        else:
            dims = (20 + rnd.randrange(4), 20, 20)
            slots, steps = 20000 + rnd.randrange(20000), rnd.randrange(5, 12)
        p = Prob(dims, slots, steps, nbytes=(1 << 20) * (1 + len(dims)))
        _FOOT[dims] = (max(_FOOT.get(dims, (0, 0))[0], slots), p.nbytes)
        probs.append(p)
    return probs


def _check(out, probs, made):
    assert sorted(r.tag.seq for r in out) == list(range(len(probs)))
    assert all(op.closed == 1 for op in made), [op.closed for op in made]


def test_serve_one_engine(fake_env):
    make_op, made = fake_env
    probs = _stream(300)
    reqs = [B.Request(0.0, k, p) for k, p in enumerate(probs)]
    bs = FakeSolver(slots=200_000, max_instances=64)
    st = {}
    out = B.serve(reqs, make_op, budget_bytes=100 * 200_000 + (64 << 20), tol=1e-6,
                  prep_workers=2, stats=st, solver=bs)
    _check(out, probs, made)
    assert all(r.status == "solved" for r in out)
    assert bs.peak <= 64 and bs.used == 0 and bs.live == 0
    assert st["loop_steps"] > 0


def test_serve_lanes_route_and_finish(fake_env):
    make_op, made = fake_env
    probs = _stream(400, seed=1)
    reqs = [B.Request(0.0, k, p) for k, p in enumerate(probs)]
    l3 = FakeSolver(slots=150_000, max_instances=8)
    l2 = FakeSolver(slots=60_000, max_instances=16)
    l1 = FakeSolver(slots=30_000, max_instances=32)
    out = B.serve(reqs, make_op, budget_bytes=100 * 240_000 + (96 << 20), tol=1e-6,
                  prep_workers=2,
                  lanes=[(l3, lambda p: len(p.dims) == 3), (l2, lambda p: len(p.dims) == 2),
                         (l1, None)])
    _check(out, probs, made)
    assert all(r.status == "solved" for r in out)
    by_uid = {p.uid: p for p in probs}
    assert all(len(by_uid[u].dims) == 3 for u in l3.served)
    assert all(len(by_uid[u].dims) == 2 for u in l2.served)
    assert all(len(by_uid[u].dims) == 1 for u in l1.served)
    assert len(l3.served) + len(l2.served) + len(l1.served) == len(probs)
    for lane, cap in ((l3, 8), (l2, 16), (l1, 32)):
        assert lane.peak <= cap and lane.used == 0 and lane.live == 0
    # the lanes' steps ran on the serve() lane threads, not the caller's
    assert threading.get_ident() not in l1.threads | l2.threads | l3.threads


def test_serve_lanes_tight_handle_budget(fake_env):
    """A handle budget that holds only a few handles at a time: requests wait
    for handle bytes and still all finish (no leaked bytes, no stall)."""
    make_op, made = fake_env
    probs = _stream(150, seed=2)
    reqs = [B.Request(0.0, k, p) for k, p in enumerate(probs)]
    l3 = FakeSolver(slots=150_000, max_instances=8)
    rest = FakeSolver(slots=90_000, max_instances=32)
    pool = 100 * (150_000 + 90_000)
    out = B.serve(reqs, make_op, budget_bytes=pool + (12 << 20), tol=1e-6, prep_workers=2,
                  lanes=[(l3, lambda p: len(p.dims) == 3), (rest, None)])
    _check(out, probs, made)
    assert all(r.status == "solved" for r in out)


def test_serve_lanes_rejections(fake_env):
    """A problem no lane takes and one larger than its lane's pool come back
    "rejected" at once; the others are served."""
    make_op, made = fake_env
    probs = _stream(60, seed=3)
    big = Prob((99, 99), 10 ** 9, 3)
    _FOOT[big.dims] = (big.slots, big.nbytes)
    odd = Prob((7, 7, 7), 100, 3)
    _FOOT[odd.dims] = (odd.slots, odd.nbytes)
    probs = probs + [big, odd]
    reqs = [B.Request(0.0, k, p) for k, p in enumerate(probs)]
    l2 = FakeSolver(slots=60_000, max_instances=16)
    l1 = FakeSolver(slots=30_000, max_instances=32)
    out = B.serve(reqs, make_op, budget_bytes=100 * 90_000 + (64 << 20), tol=1e-6,
                  prep_workers=2,
                  lanes=[(l2, lambda p: len(p.dims) == 2),
                         (l1, lambda p: len(p.dims) == 1)])
    assert sorted(r.tag.seq for r in out) == list(range(len(probs)))
    st = {r.tag.seq: r for r in out}
    assert st[len(probs) - 2].status == "rejected" and "pool" in st[len(probs) - 2].error
    assert st[len(probs) - 1].status == "rejected" and "no lane" in st[len(probs) - 1].error
    three = [k for k, p in enumerate(probs[:-2]) if len(p.dims) == 3]
    assert all(st[k].status == "rejected" for k in three)
    assert all(st[k].status == "solved" for k in range(len(probs) - 2) if k not in three)
    assert all(op.closed == 1 for op in made)


def test_serve_rejects_solver_and_lanes(fake_env):
    make_op, _ = fake_env
    with pytest.raises(ValueError):
        B.serve([], make_op, solver=FakeSolver(10, 1), lanes=[(FakeSolver(10, 1), None)])
