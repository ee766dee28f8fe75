"""Serving-loop scheduler on the CPU: admission under the HBM budget, first-come order,
batches of one solver class, every request served once (stand-in executor, no GPU)."""
import threading
import time

import numpy as np
import pytest

from pyoomph_amd.serving import (Request, ServingLoop, poisson_stream, size_for_dofs, summarize,
                                 _NDOF)


class FakeExecutor:
    """Sleeps in proportion to the batch's unknowns; records what ran together."""

    def __init__(self, per_dof=2e-9):
        self.per_dof = per_dof
        self.lock = threading.Lock()
        self.running = 0
        self.max_running = 0
        self.batches = []

    def make_stream(self):
        return None

    def run_batch(self, batch, stream):
        with self.lock:
            self.running += 1
            self.max_running = max(self.max_running, self.running)
            self.batches.append([r.rid for r in batch])
        time.sleep(0.001 + self.per_dof * sum(r.ndof for r in batch))
        with self.lock:
            self.running -= 1
        return [(np.full(3, float(r.rid)), {"converged": True}) for r in batch]


def test_sizes_and_stream_are_seeded():
    a = poisson_stream(500, 40, seed=3)
    b = poisson_stream(500, 40, seed=3)
    assert [(r.kind, r.N, r.seed, r.arrival) for r in a] == [(r.kind, r.N, r.seed, r.arrival) for r in b]
    assert all(x.arrival < y.arrival for x, y in zip(a, a[1:]))
    for r in a:
        assert r.ndof >= 5_000 * 0.999
        lim = 12_000 if r.kind == "ns" else 600_000
        # the smallest size with at least the drawn count of unknowns
        assert _NDOF[r.kind](r.N - 1) < lim
    for kind in ("poisson", "rd", "ns"):
        N = size_for_dofs(kind, 10_000)
        assert _NDOF[kind](N) >= 10_000 > _NDOF[kind](N - 1)


def test_every_request_served_once_with_class_pure_batches():
    reqs = poisson_stream(2000, 60, seed=1)
    ex = FakeExecutor()
    loop = ServingLoop(ex, hbm_budget=1e12, workers=3, max_batch_dofs=2_000_000)
    res = loop.run(reqs)
    assert [r.rid for r in res] == [r.rid for r in sorted(reqs, key=lambda r: r.rid)]
    assert all(float(r.dofs[0]) == r.rid for r in res)
    seen = [rid for b in ex.batches for rid in b]
    assert sorted(seen) == list(range(len(reqs)))
    kinds = {r.rid: r.solver for r in reqs}
    for b in ex.batches:
        assert len({kinds[i] for i in b}) == 1
        assert sum(reqs[i].ndof for i in b) <= 2_000_000 or len(b) == 1
    assert ex.max_running <= 3
    for r in res:
        assert r.arrival <= r.admitted + 1e-6 and r.admitted <= r.started + 1e-6 and r.started <= r.done
    s = summarize(res)
    assert s["requests"] == len(reqs) and s["p99_s"] >= s["p50_s"] > 0


def test_budget_is_never_exceeded_and_admission_is_fifo():
    reqs = [Request(i, "poisson" if i % 2 else "rd", 20, i, 0.0) for i in range(24)]
    est = lambda r: 1000.0 if r.kind == "rd" else 400.0
    ex = FakeExecutor(per_dof=1e-8)
    loop = ServingLoop(ex, hbm_budget=2500.0, workers=4, max_batch_dofs=10**9, estimator=est)
    res = loop.run(reqs, realtime=False)
    assert loop.peak_reserved <= 2500.0
    adm = sorted(res, key=lambda r: r.admitted)
    assert [r.rid for r in adm] == sorted(r.rid for r in res)
    assert len(res) == 24


def test_request_larger_than_budget_is_rejected():
    loop = ServingLoop(FakeExecutor(), hbm_budget=10.0)
    with pytest.raises(ValueError):
        loop.run([Request(0, "rd", 30, 0, 0.0)])


def test_executor_failure_surfaces():
    class Bad(FakeExecutor):
        def run_batch(self, batch, stream):
            raise RuntimeError("device failure")
    with pytest.raises(RuntimeError):
        ServingLoop(Bad(), workers=2).run(poisson_stream(1000, 5, seed=0))
