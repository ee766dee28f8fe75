"""Serving loop: Newton solves of different meshes and parameters as they arrive.

pyoomph has no request loop; its closest analogue is the process farm of parameter scans
(pyoomph/utils/paramscan.py:249-277), which runs one independent Problem per process.
Here requests share one GPU:

* Arrivals. A request names a problem (its kind, mesh size and seed) and an arrival
  time; ``poisson_stream`` draws a seeded Poisson-arrival stream mixing 2D P2 Poisson,
  3D Q2 reaction-diffusion and 2D Taylor-Hood Navier-Stokes requests.
* Admission. A request is admitted, first come first served, when its estimated device
  footprint fits in what is left of the HBM budget; the head of the queue is never
  overtaken, so no request starves.
* Continuous batching. Admitted requests wait in one queue per solver class (conjugate
  gradients for the symmetric forms, GMRES otherwise). Whenever one of the ``workers``
  is free it takes the oldest admitted request and every admitted request of the same
  class behind it, up to ``max_batch_dofs``, builds them into one batched problem on its
  own HIP stream and solves it. Requests that arrive while batches run are admitted as
  memory frees up and start as soon as a worker is free, so batches of different
  arrivals overlap on the device.
* Each sub-problem of a batch has its own Newton and Krylov scalars and reductions, so a
  request's result does not depend on the batch it ran in (``tests/test_gpu_serving.py``
  checks alone against in the mix).

The executor is pluggable: ``GpuExecutor`` runs batches with the HIP library;
the scheduler itself is plain host code and is tested with a stand-in executor.
"""
import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

__all__ = ["Request", "Result", "poisson_stream", "ServingLoop", "GpuExecutor", "estimate_bytes",
           "SOLVER_CLASS"]

# solver class of each request kind (symmetric positive definite Jacobian -> CG)
SOLVER_CLASS = {"poisson": "cg", "rd": "cg", "ns": "gmres"}


def _ndof_poisson(N):
    return (2 * N - 1) ** 2


def _ndof_rd(N):
    return (2 * N - 1) ** 3


def _ndof_ns(N):
    # Q2 velocities on (2N+1)^2 nodes minus the boundary, Q1 pressure on (N+1)^2 minus one
    return 2 * (2 * N - 1) ** 2 + (N + 1) ** 2 - 1


_NDOF = {"poisson": _ndof_poisson, "rd": _ndof_rd, "ns": _ndof_ns}


def size_for_dofs(kind, ndof):
    """Smallest mesh size N of a request kind with at least ``ndof`` unknowns."""
    N = 2
    while _NDOF[kind](N) < ndof:
        N += 1
    return N


# device bytes per unknown of one request, measured on the library's layout (pattern,
# element buffers, Krylov workspace) and rounded up; used only for admission
_BYTES_PER_DOF = {"poisson": 1400, "rd": 5200, "ns": 4200}


def estimate_bytes(kind, ndof):
    return int(_BYTES_PER_DOF[kind] * ndof) + (1 << 20)


@dataclass
class Request:
    rid: int
    kind: str          # "poisson" | "rd" | "ns"
    N: int             # mesh size (elements per side)
    seed: int
    arrival: float     # seconds after the start of the stream
    payload: object = field(default=None, repr=False, compare=False)   # the request's mesh problem

    @property
    def ndof(self):
        return _NDOF[self.kind](self.N)

    @property
    def solver(self):
        return SOLVER_CLASS[self.kind]

    def attach_payload(self):
        """Builds the request's mesh problem now (the client side of a request: it
        arrives with its mesh, fields and parameters)."""
        self.payload = self.mesh_problem()
        return self

    def mesh_problem(self):
        if self.payload is not None:
            return self.payload
        from . import workloads as wl
        if self.kind == "poisson":
            return wl.poisson_gaussians(self.N, seed=self.seed)
        if self.kind == "rd":
            return wl.reaction_diffusion_cube(self.N, seed=self.seed)
        return wl.lid_driven_cavity(self.N, Re=100.0)


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    started: float
    done: float
    batch: int
    batch_size: int
    dofs: object = field(repr=False, default=None)
    stats: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.done - self.arrival


def poisson_stream(rate, n, seed=0, mix=(("poisson", 0.45), ("rd", 0.35), ("ns", 0.20)),
                   dof_range=(5_000, 600_000), ns_max_dofs=12_000):
    """``n`` requests with exponential inter-arrival times of mean 1/rate (seeded). Sizes
    are log-uniform in ``dof_range``; Navier-Stokes requests are capped at
    ``ns_max_dofs`` unknowns (the GMRES/ILU(0) solve of larger Taylor-Hood systems is
    slow)."""
    rng = np.random.default_rng(seed)
    kinds = [k for k, _ in mix]
    p = np.array([w for _, w in mix], dtype=float)
    p /= p.sum()
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        kind = kinds[int(rng.choice(len(kinds), p=p))]
        lo, hi = dof_range
        if kind == "ns":
            hi = min(hi, ns_max_dofs)
        nd = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        out.append(Request(i, kind, size_for_dofs(kind, nd), int(rng.integers(1 << 30)), t))
    return out


class ServingLoop:
    """Admission by HBM budget, continuous batching per solver class, ``workers`` batches
    in flight (one HIP stream each)."""

    def __init__(self, executor, hbm_budget=200e9, workers=4, max_batch_dofs=4_000_000,
                 estimator=None):
        self.executor = executor
        self.budget = float(hbm_budget)
        self.workers = int(workers)
        self.max_batch_dofs = int(max_batch_dofs)
        self.estimator = estimator or (lambda r: estimate_bytes(r.kind, r.ndof))
        self.peak_reserved = 0.0
        self.batches = []        # (batch id, [rid...], solver)

    def run(self, requests, realtime=True):
        """Serve ``requests`` (sorted by arrival). With ``realtime`` the arrivals are
        released at their times on the wall clock; otherwise all are present at t = 0.
        Returns the results in request order."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            if self.estimator(r) > self.budget:
                raise ValueError("request %d needs more than the whole HBM budget" % r.rid)
        cv = threading.Condition()
        state = {"reserved": 0.0, "idle": self.workers, "next": 0, "error": None}
        pending, queues, results = [], {}, {}
        admitted_at = {}
        work = []                 # batches handed to workers
        stop = [False]
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        def worker(wid):
            stream = self.executor.make_stream()
            while True:
                with cv:
                    while not work and not stop[0]:
                        cv.wait()
                    if not work:
                        return
                    bid, batch = work.pop(0)
                t_start = now()
                try:
                    outs = self.executor.run_batch(batch, stream)
                except Exception as e:   # surfaced to run()
                    with cv:
                        state["error"] = e
                        cv.notify_all()
                    return
                t_done = now()
                with cv:
                    for r, (x, st) in zip(batch, outs):
                        results[r.rid] = Result(r.rid, r.arrival, admitted_at[r.rid], t_start, t_done,
                                                bid, len(batch), x, st)
                        state["reserved"] -= self.estimator(r)
                    state["idle"] += 1
                    cv.notify_all()

        threads = [threading.Thread(target=worker, args=(w,), daemon=True) for w in range(self.workers)]
        for th in threads:
            th.start()
        nb = 0
        try:
            with cv:
                while len(results) < len(reqs):
                    if state["error"] is not None:
                        raise state["error"]
                    t = now()
                    # arrivals
                    while state["next"] < len(reqs) and (not realtime or reqs[state["next"]].arrival <= t):
                        pending.append(reqs[state["next"]])
                        state["next"] += 1
                    # admission: first come first served under the budget
                    while pending and state["reserved"] + self.estimator(pending[0]) <= self.budget:
                        r = pending.pop(0)
                        state["reserved"] += self.estimator(r)
                        self.peak_reserved = max(self.peak_reserved, state["reserved"])
                        admitted_at[r.rid] = now()
                        queues.setdefault(r.solver, []).append(r)
                    # batching: the oldest admitted request and its class mates
                    while state["idle"] > 0 and any(queues.values()):
                        cls = min((q[0].arrival, c) for c, q in queues.items() if q)[1]
                        q = queues[cls]
                        batch = [q.pop(0)]
                        dofs = batch[0].ndof
                        while q and dofs + q[0].ndof <= self.max_batch_dofs:
                            dofs += q[0].ndof
                            batch.append(q.pop(0))
                        work.append((nb, batch))
                        self.batches.append((nb, [r.rid for r in batch], cls))
                        nb += 1
                        state["idle"] -= 1
                        cv.notify_all()
                    if len(results) == len(reqs):
                        break
                    timeout = None
                    if state["next"] < len(reqs) and realtime:
                        timeout = max(0.0, reqs[state["next"]].arrival - now())
                    cv.wait(timeout=timeout if timeout is not None else 0.5)
        finally:
            with cv:
                stop[0] = True
                cv.notify_all()
            for th in threads:
                th.join(timeout=60)
        return [results[r.rid] for r in sorted(reqs, key=lambda r: r.rid)]


def summarize(results):
    """solves/s over the stream and latency percentiles (seconds)."""
    lat = np.array([r.latency for r in results])
    first = min(r.arrival for r in results)
    last = max(r.done for r in results)
    return {"requests": len(results), "solves_per_s": len(results) / max(last - first, 1e-9),
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_batch": float(np.mean([r.batch_size for r in results]))}


class GpuExecutor:
    """Runs a batch with the HIP library: one batched ``Problem`` on the worker's stream,
    FP64 Newton; CG for the symmetric class (Jacobi or the multigrid V-cycle,
    ``cg_precond``), GMRES(``gmres_restart``) + ILU(0) otherwise."""

    def __init__(self, device="cuda:0", cg_precond="jacobi", gmres_restart=200):
        import torch
        self.torch = torch
        self.device = torch.device(device)
        self.cg_precond = cg_precond
        self.gmres_restart = gmres_restart

    def make_stream(self):
        return self.torch.cuda.Stream(device=self.device)

    def configure(self, P, cls):
        if cls == "cg":
            P.krylov, P.preconditioner = "cg", self.cg_precond
        else:
            P.krylov, P.preconditioner = "gmres", "ilu0"
            P.gmres_restart = self.gmres_restart
            P.newton_solver_tolerance = 1e-10
        P.gmres_max_iters = 20000

    def run_batch(self, batch, stream):
        from .problem import Problem
        torch = self.torch
        with torch.cuda.stream(stream):
            P = Problem([r.mesh_problem() for r in batch], device=self.device)
            self.configure(P, batch[0].solver)
            st = P.solve(raise_on_failure=False)
            x = P.get_current_dofs()
            outs = [(P.sub_dofs(i, x).copy(), dict(st[i], device_bytes=P.device_bytes))
                    for i in range(len(batch))]
            del P
        return outs

    def run_alone(self, request):
        """The same request solved by itself (for the alone-vs-mix check)."""
        return self.run_batch([request], self.torch.cuda.current_stream(self.device))[0]
