"""Serving-loop scheduler on the CPU with a stand-in executor (no GPU): every request is
served once by its mesh class's pool; slots are never over-filled; requests join a running
pool at step boundaries in arrival order and leave as soon as their solve is done; at most
``workers`` pools step at once; pools are created within the HBM budget; a failing step
surfaces."""
import threading
import time

import numpy as np
import pytest

from pyoomph_amd.serving import SIZES, Request, ServingLoop, poisson_stream, slots_for, summarize, _NDOF


class FakePool:
    """A request needs 1 + seed % 4 Newton steps; a step sleeps in proportion to the busy
    unknowns."""

    def __init__(self, ex, sig, slots):
        self.ex, self.sig, self.slots = ex, sig, slots
        self.device_bytes = ex.bytes_per_slot * slots
        self.busy = {}
        self.stream = None
        self.log = []            # (step index, frozenset of rids in the step)
        self.loads = []          # rids in load order

    def bind(self, stream):
        self.stream = stream

    def load(self, slot, r):
        assert slot not in self.busy and 0 <= slot < self.slots
        self.busy[slot] = [r, 0]
        self.loads.append(r.rid)

    def step(self):
        ex = self.ex
        with ex.lock:
            ex.running += 1
            ex.max_running = max(ex.max_running, ex.running)
            # no two pools step on one stream at the same time
            assert self.stream is not None and self.stream not in ex.streams_in_use
            ex.streams_in_use.add(self.stream)
        self.log.append(frozenset(r.rid for r, _ in self.busy.values()))
        if ex.fail_after is not None and sum(len(p.log) for p in ex.pools) > ex.fail_after:
            with ex.lock:
                ex.streams_in_use.discard(self.stream)
            raise RuntimeError("device failure")
        time.sleep(0.0005 + ex.per_dof * sum(r.ndof for r, _ in self.busy.values()))
        done = []
        for slot, st in self.busy.items():
            st[1] += 1
            if st[1] >= 1 + st[0].seed % 4:
                done.append(slot)
        with ex.lock:
            ex.running -= 1
            ex.streams_in_use.discard(self.stream)
        return done

    def take(self, slot):
        r, n = self.busy.pop(slot)
        return np.full(3, float(r.rid)), {"newton_iters": n, "converged": True}

    def close(self):
        pass


class FakeExecutor:
    def __init__(self, per_dof=2e-9, bytes_per_slot=1000, fail_after=None):
        self.per_dof, self.bytes_per_slot, self.fail_after = per_dof, bytes_per_slot, fail_after
        self.lock = threading.Lock()
        self.running = self.max_running = 0
        self.pools = []
        self.streams_made = 0
        self.streams_in_use = set()

    def make_stream(self):
        self.streams_made += 1
        return "stream%d" % self.streams_made

    def create(self, sig, slots):
        p = FakePool(self, sig, slots)
        self.pools.append(p)
        return p


def all_sigs():
    return [(k, N) for k, v in SIZES.items() for N in v]


def test_stream_is_seeded_and_sizes_are_the_served_ones():
    a = poisson_stream(500, 40, seed=3)
    b = poisson_stream(500, 40, seed=3)
    assert [(r.kind, r.N, r.seed, r.arrival) for r in a] == [(r.kind, r.N, r.seed, r.arrival) for r in b]
    assert all(x.arrival < y.arrival for x, y in zip(a, a[1:]))
    assert all(r.N in SIZES[r.kind] for r in a)
    for k, sizes in SIZES.items():
        nd = [_NDOF[k](N) for N in sizes]
        assert nd == sorted(nd) and nd[0] >= 5_000 and nd[-1] <= 600_000
    assert slots_for("poisson", 36) == 16 and slots_for("rd", 40) == 8


def test_every_request_served_once_by_its_pool_in_arrival_order():
    reqs = poisson_stream(3000, 200, seed=1)
    ex = FakeExecutor()
    loop = ServingLoop(ex, hbm_budget=1e12, workers=3, slots=lambda k, N: 3)
    loop.prepare(all_sigs())
    res = loop.run(reqs)
    assert [r.rid for r in res] == list(range(len(reqs)))
    assert all(float(r.dofs[0]) == r.rid for r in res)
    assert all(r.pool == q.signature for r, q in zip(res, sorted(reqs, key=lambda q: q.rid)))
    assert ex.max_running <= 3
    by_rid = {q.rid: q for q in reqs}
    for p in ex.pools:
        # never more busy slots than the pool has; FIFO within the pool
        assert all(len(s) <= p.slots for s in p.log)
        assert p.loads == sorted(p.loads, key=lambda i: by_rid[i].arrival)
        # each request stays for exactly its step count, in consecutive steps of the pool
        for rid in p.loads:
            idx = [i for i, s in enumerate(p.log) if rid in s]
            assert idx == list(range(idx[0], idx[0] + 1 + by_rid[rid].seed % 4))
    for r in res:
        assert r.arrival <= r.admitted + 1e-6 and r.admitted <= r.done
        assert r.stats["newton_iters"] == 1 + by_rid[r.rid].seed % 4
    s = summarize(res)
    assert s["requests"] == len(reqs) and s["p99_s"] >= s["p50_s"] > 0


def test_newcomers_join_a_running_pool_and_finished_ones_leave():
    # one pool, 4 slots; a long request (4 steps) and short ones arriving later
    reqs = [Request(0, "poisson", 36, 3, 0.0)] + [Request(i, "poisson", 36, 4 * i, 0.002 * i) for i in range(1, 9)]
    ex = FakeExecutor(per_dof=1e-6)
    loop = ServingLoop(ex, hbm_budget=1e12, workers=1, slots=lambda k, N: 4)
    loop.prepare([("poisson", 36)])
    res = loop.run(reqs)
    (p,) = ex.pools
    # some step ran the long request together with a request admitted after it started
    first = [i for i, s in enumerate(p.log) if 0 in s]
    assert any(len(p.log[i] - {0}) > 0 and any(x not in p.log[first[0]] for x in p.log[i]) for i in first[1:])
    # the short requests (1 step each) finished before the long one
    assert all(r.done <= res[0].done for r in res[1:3])


def test_pools_are_created_within_the_budget():
    loop = ServingLoop(FakeExecutor(bytes_per_slot=100), hbm_budget=1000.0, slots=lambda k, N: 4)
    loop.prepare([("poisson", 36), ("rd", 10)])
    assert loop.reserved == 800
    with pytest.raises(MemoryError):
        loop.prepare([("ns", 26)])
    assert loop.reserved == 800 and ("ns", 26) not in loop.pool_of


def test_request_without_a_pool_is_rejected():
    loop = ServingLoop(FakeExecutor())
    loop.prepare([("poisson", 36)])
    with pytest.raises(ValueError):
        loop.run([Request(0, "rd", 10, 0, 0.0)])


def test_step_failure_surfaces():
    loop = ServingLoop(FakeExecutor(fail_after=5), workers=2, slots=lambda k, N: 2)
    loop.prepare(all_sigs())
    with pytest.raises(RuntimeError, match="device failure"):
        loop.run(poisson_stream(1000, 30, seed=0))
