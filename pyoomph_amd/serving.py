"""Serving loop: Newton solves of different meshes and parameters as they arrive.

pyoomph has no request loop; its closest analogue is the process farm of parameter scans
(pyoomph/utils/paramscan.py:249-277), which runs one independent Problem per process.
Here requests share one GPU, batched continuously at the granularity of a Newton step:

* Mesh classes. A request names a problem kind (2D P2 Poisson, 3D Q2 reaction-diffusion,
  2D Taylor-Hood Navier-Stokes), one of the kind's mesh sizes (``SIZES``: 5k to 490k
  unknowns) and its own data: the Poisson source's eight Gaussians, the
  reaction-diffusion source field, the Navier-Stokes Reynolds number, all drawn from the
  request's seed. ``poisson_stream`` draws a seeded Poisson-arrival stream of them.
* Slot pools. Every (kind, mesh size) has a pool: one device-resident batched problem
  whose sub-problems are ``slots`` copies of that mesh. Mesh, equation maps, sparsity
  pattern, solver plans (ILU order, multigrid aggregates) and Krylov workspace are built
  once, when the server starts, and kept; pools are created while their measured device
  bytes fit in the HBM budget. A request occupies one slot: its parameters / source
  field are copied into the slot and the slot's dofs set to the start value.
* Continuous batching. Each pool advances its busy slots one Newton step at a time
  (``pyhip_newton_step``). Requests that arrive while a pool is stepping wait in the
  pool's queue and take free slots at the next step boundary, first come first served;
  a request whose solve has converged (or failed) leaves at once -- its dofs copied to
  the host, its slot freed -- while the others go on. At most ``workers`` pools step at
  a time, each on one of ``workers`` HIP streams (one hardware queue each), so different
  pools' work overlaps on the device; pools wanting to step are served in the order they
  asked.
* Each slot's Newton and Krylov iteration is its own (per-sub-problem element rows,
  scalars, reductions and convergence tests; Jacobi per row; the multigrid hierarchies,
  per sub-problem, rebuilt from every step's Jacobian), so a request gives the same
  bits alone and in the mix (``tests/test_gpu_serving.py``).

The executor is pluggable: ``GpuPools`` runs pools with the HIP library; the scheduler is
plain host code, tested with a stand-in executor (``tests/test_serving_scheduler.py``).
"""
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

__all__ = ["Request", "Result", "poisson_stream", "attach_payloads", "request_data", "ServingLoop",
           "GpuPools", "SlotPool", "summarize", "slots_for", "SIZES", "SOLVER_CLASS"]

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

# the mesh sizes served per kind, about log-uniform in unknowns: Poisson 5k..450k, reaction-
# diffusion 7k..490k, Taylor-Hood 6k..360k
SIZES = {"poisson": (36, 56, 88, 136, 212, 336),
         "rd": (10, 14, 20, 28, 40),
         "ns": (26, 44, 76, 128, 200)}


def slots_for(kind, N, target_dofs=4_000_000, lo=2, hi=16):
    """Slots of a pool: about ``target_dofs`` unknowns per pool, within [lo, hi]."""
    return int(min(hi, max(lo, target_dofs // _NDOF[kind](N))))


def ns_reynolds(seed):
    """Reynolds number of a Navier-Stokes request: 50..150 from its seed."""
    return 50.0 + float(seed % 101)


def request_data(kind, N, seed, mesh=None):
    """(parameters, auxiliary nodal field or None) of a request; ``mesh`` is the request's
    mesh (needed for the reaction-diffusion field)."""
    from . import workloads as wl
    from .equations import NavierStokesEquations, PoissonEquation
    if kind == "poisson":
        return PoissonEquation("u", space="C2", source=wl.gaussian_source(seed)).params(), None
    if kind == "rd":
        if mesh is None:
            from .meshes import CuboidBrickMesh
            mesh = CuboidBrickMesh(N)
        return np.zeros(0), wl.smoothed_noise(mesh, seed)
    if kind == "ns":
        return NavierStokesEquations(dynamic_viscosity=1.0 / ns_reynolds(seed), mass_density=1.0).params(), None
    raise KeyError(kind)


@dataclass
class Request:
    rid: int
    kind: str          # "poisson" | "rd" | "ns"
    N: int             # mesh size (elements per side)
    seed: int          # the request's data
    arrival: float     # seconds after the start of the stream
    payload: object = field(default=None, repr=False, compare=False)   # (params, aux)

    @property
    def ndof(self):
        return _NDOF[self.kind](self.N)

    @property
    def solver(self):
        return SOLVER_CLASS[self.kind]

    @property
    def signature(self):
        return (self.kind, self.N)

    def data(self, mesh=None):
        if self.payload is None:
            self.payload = request_data(self.kind, self.N, self.seed, mesh)
        return self.payload

    def mesh_problem(self):
        """The request as a stand-alone mesh problem (its own mesh and data)."""
        from . import workloads as wl
        if self.kind == "poisson":
            return wl.poisson_gaussians(self.N, seed=self.seed)
        if self.kind == "rd":
            return wl.reaction_diffusion_cube(self.N, seed=self.seed)
        return wl.lid_driven_cavity(self.N, Re=ns_reynolds(self.seed))


def attach_payloads(requests):
    """Computes every request's data (the client side: a request arrives with its fields
    and parameters), one mesh per mesh class."""
    meshes = {}
    for r in requests:
        if r.payload is not None:
            continue
        if r.kind == "rd" and r.signature not in meshes:
            from .meshes import CuboidBrickMesh
            meshes[r.signature] = CuboidBrickMesh(r.N)
        r.data(meshes.get(r.signature))
    return requests


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float    # loaded into a slot
    done: float        # converged and copied to the host
    pool: tuple
    slot: int
    dofs: object = field(repr=False, default=None)
    stats: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.done - self.arrival


def poisson_stream(rate, n, seed=0, mix=(("poisson", 0.45), ("rd", 0.35), ("ns", 0.20)), sizes=None):
    """``n`` requests with exponential inter-arrival times of mean 1/rate (seeded); kinds
    drawn with the ``mix`` weights, mesh sizes uniformly from the kind's ``sizes``."""
    sizes = sizes or SIZES
    rng = np.random.default_rng(seed)
    kinds = [k for k, _ in mix]
    p = np.array([w for _, w in mix], dtype=float)
    p /= p.sum()
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        kind = kinds[int(rng.choice(len(kinds), p=p))]
        N = int(sizes[kind][int(rng.integers(len(sizes[kind])))])
        out.append(Request(i, kind, N, int(rng.integers(1 << 30)), t))
    return out


class ServingLoop:
    """Slot pools per mesh class, Newton-step-granular continuous batching, at most
    ``workers`` pools stepping at a time.

    ``executor``: ``create(signature, slots) -> pool`` and ``make_stream()``; a pool has
    ``slots``, ``device_bytes``, ``bind(stream)``, ``load(slot, request)``,
    ``step() -> [finished slots]`` and ``take(slot) -> (dofs, stats)``."""

    def __init__(self, executor, hbm_budget=200e9, workers=4, slots=None):
        self.executor = executor
        self.budget = float(hbm_budget)
        self.workers = int(workers)
        self.slots = slots or slots_for
        self.reserved = 0
        self.pool_of = {}
        self.steps = {}          # signature -> [busy slots per step]

    def prepare(self, signatures):
        """Creates the pools of these mesh classes (server start, before any request)."""
        for sig in signatures:
            sig = tuple(sig)
            if sig in self.pool_of:
                continue
            pool = self.executor.create(sig, self.slots(*sig))
            if self.reserved + pool.device_bytes > self.budget:
                pool.close()
                raise MemoryError("the pool of %s (%d bytes) does not fit in the HBM budget" % (sig, pool.device_bytes))
            self.reserved += pool.device_bytes
            self.pool_of[sig] = pool
            self.steps[sig] = []

    def close(self):
        for pool in self.pool_of.values():
            pool.close()
        self.pool_of.clear()
        self.reserved = 0

    def run(self, requests, realtime=True):
        """Serves ``requests``; with ``realtime`` each arrives at its time on the wall
        clock, otherwise all at t = 0. Returns the results in request order."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        missing = {r.signature for r in reqs} - set(self.pool_of)
        if missing:
            raise ValueError("no pool serves the mesh classes %s" % sorted(missing))
        cv = threading.Condition()
        queue = {sig: deque() for sig in self.pool_of}
        results = {}
        turn = deque()                     # pools waiting to step, in the order they asked
        state = {"released": 0, "tokens": self.workers, "error": None, "closing": False}
        # one stream per step in flight: a stepping pool runs on a free one (more streams
        # than hardware queues would put pools behind each other's kernels)
        streams = [self.executor.make_stream() for _ in range(self.workers)]
        free_streams = deque(range(self.workers))
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        def pool_worker(sig):
            pool = self.pool_of[sig]
            busy = {}                      # slot -> (request, admitted)
            free = deque(range(pool.slots))
            while True:
                with cv:
                    while not busy and not queue[sig] and not state["closing"]:
                        cv.wait()
                    if not busy and not queue[sig]:
                        return
                    turn.append(sig)
                    while (turn[0] != sig or state["tokens"] == 0) and state["error"] is None:
                        cv.wait()
                    if state["error"] is not None:
                        return
                    turn.popleft()
                    state["tokens"] -= 1
                    si = free_streams.popleft()
                    cv.notify_all()
                    admit = []
                    while free and queue[sig]:
                        admit.append((free.popleft(), queue[sig].popleft()))
                try:
                    pool.bind(streams[si])
                    t_adm = now()
                    for slot, r in admit:
                        pool.load(slot, r)
                        busy[slot] = (r, t_adm)
                    t_s = time.perf_counter()
                    finished = pool.step()
                    self.steps[sig].append((len(busy), time.perf_counter() - t_s))
                    outs = [(slot, pool.take(slot)) for slot in finished]
                except BaseException as e:   # surfaced by run()
                    with cv:
                        state["error"] = e
                        state["tokens"] += 1
                        free_streams.append(si)
                        cv.notify_all()
                    return
                t_done = now()
                with cv:
                    state["tokens"] += 1
                    free_streams.append(si)
                    for slot, (x, st) in outs:
                        r, t_a = busy.pop(slot)
                        free.append(slot)
                        results[r.rid] = Result(r.rid, r.arrival, t_a, t_done, sig, slot, x, st)
                    cv.notify_all()

        threads = [threading.Thread(target=pool_worker, args=(sig,), daemon=True) for sig in self.pool_of]
        for th in threads:
            th.start()
        try:
            with cv:
                while len(results) < len(reqs):
                    if state["error"] is not None:
                        raise state["error"]
                    t = now()
                    while state["released"] < len(reqs) and (not realtime or reqs[state["released"]].arrival <= t):
                        r = reqs[state["released"]]
                        queue[r.signature].append(r)
                        state["released"] += 1
                        cv.notify_all()
                    timeout = 0.5
                    if state["released"] < len(reqs) and realtime:
                        timeout = min(timeout, max(0.0, reqs[state["released"]].arrival - now()))
                    cv.wait(timeout=timeout)
                if state["error"] is not None:
                    raise state["error"]
        finally:
            with cv:
                state["closing"] = True
                if state["error"] is None and len(results) < len(reqs):
                    state["error"] = RuntimeError("serving stopped")
                cv.notify_all()
            for th in threads:
                th.join(timeout=120)
        return [results[r.rid] for r in sorted(reqs, key=lambda r: r.rid)]

    def occupancy(self):
        """Per pool: Newton steps taken, mean busy slots per step, mean wall time of a step
        (load, step and take of the finished slots)."""
        out = {}
        for sig, v in self.steps.items():
            b = [x[0] for x in v]
            t = [x[1] for x in v]
            out["%s-%d" % sig] = {"slots": self.pool_of[sig].slots if sig in self.pool_of else None,
                                  "steps": len(v), "mean_busy": float(np.mean(b)) if b else 0.0,
                                  "ms_per_step": 1e3 * float(np.mean(t)) if t else 0.0}
        return out


def summarize(results):
    """solves/s over the stream and latency percentiles (seconds)."""
    lat = np.array([r.latency for r in results])
    first = min(r.arrival for r in results)
    last = max(r.done for r in results)
    return {"requests": len(results), "solves_per_s": len(results) / max(last - first, 1e-9),
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_newton_steps": float(np.mean([r.stats.get("newton_iters", 0) for r in results])),
            "converged": int(sum(bool(r.stats.get("converged", False)) for r in results))}


class SlotPool:
    """One mesh class's device-resident batched problem (``slots`` copies of the mesh) on
    its own HIP stream."""

    def __init__(self, executor, sig, slots, warm=True):
        from . import _lib
        from .problem import Problem
        torch = executor.torch
        self.executor, self.sig, self.slots = executor, tuple(sig), int(slots)
        kind, N = self.sig
        self.stream = torch.cuda.Stream(device=executor.device)
        proto = Request(-1, kind, N, 0, 0.0)
        mp = proto.mesh_problem()
        self.mesh = mp.mesh
        with torch.cuda.stream(self.stream):
            self.P = Problem([mp] * self.slots, device=executor.device)
        executor.configure(self.P, kind)
        self.opts = self.P.newton_options()
        self._lib = _lib
        self.state = (_lib.NewtonStepState * self.slots)()
        self.mask = np.zeros(self.slots, dtype=np.int32)
        if warm:
            # one solve in slot 0 allocates the solver workspace and builds its plans
            self.load(0, proto)
            while not self.step():
                pass
            self.take(0)
        self.device_bytes = self.P.current_device_bytes()

    def bind(self, stream):
        """Runs the pool's next calls on ``stream`` (a stream of the serving loop)."""
        self.P.set_stream(stream)

    def load(self, slot, r):
        P = self.P
        params, aux = r.data(self.mesh)
        P.set_sub_data(slot, params if len(params) else None, aux)
        a, b = int(P.dof_offsets[slot]), int(P.dof_offsets[slot + 1])
        with self.executor.torch.cuda.stream(P._stream):
            P.dofs[a:b].zero_()
        self.state[slot] = self._lib.NewtonStepState()
        self.mask[slot] = 1

    def step(self):
        self.P.newton_step(self.mask, self.state, self.opts)
        return [s for s in range(self.slots) if self.mask[s] and (self.state[s].converged or self.state[s].failed)]

    def take(self, slot):
        P = self.P
        a, b = int(P.dof_offsets[slot]), int(P.dof_offsets[slot + 1])
        x = P._host(P.dofs[a:b]).copy()
        s = self.state[slot]
        self.mask[slot] = 0
        return x, {"newton_iters": s.newton_iters, "converged": bool(s.converged), "final_maxres": s.maxres,
                   "max_update": s.max_update, "linear_iters": s.linear_iters,
                   "linear_failures": s.linear_failures}

    def close(self):
        self.P = None


# CG preconditioner per symmetric kind: the multigrid V-cycle (whose first coarse level is
# the linear functions on the element vertices) for the P2 Poisson requests and the Q2
# bricks; an 8-slot pool of 490 k-unknown bricks solves in 267 ms against 705 ms with
# Jacobi, whose iterations grow with the mesh (aggregation alone: 2.3 s, its Galerkin
# products on the brick stencil cost more than the iterations save)
DEFAULT_CG_PRECOND = {"poisson": "amg", "rd": "amg"}


class GpuPools:
    """Slot pools on the GPU, FP64: CG for the symmetric kinds (``cg_precond``: "jacobi",
    "amg", or a dict per kind), GMRES(``gmres_restart``) with the block-triangular
    preconditioner (multigrid on the velocity components, lumped pressure mass for the
    Schur complement) for Taylor-Hood. Inexact
    Newton: each linear solve to ``krylov_rtol``, Newton done once max|R| is below
    ``tol`` (``ns_tol`` for Taylor-Hood) and the last update below ``correction_tol`` of
    the solution."""

    def __init__(self, device="cuda:0", cg_precond=None, gmres_restart=100, krylov_rtol=1e-4,
                 correction_tol=1e-7, tol=1e-8, ns_tol=1e-10, max_newton=30):
        import torch
        self.torch = torch
        self.device = torch.device(device)
        cg_precond = cg_precond or DEFAULT_CG_PRECOND
        self.cg_precond = dict(cg_precond) if isinstance(cg_precond, dict) else \
            {k: cg_precond for k, c in SOLVER_CLASS.items() if c == "cg"}
        self.gmres_restart = gmres_restart
        self.krylov_rtol, self.correction_tol = krylov_rtol, correction_tol
        self.tol, self.ns_tol, self.max_newton = tol, ns_tol, max_newton

    def make_stream(self):
        return self.torch.cuda.Stream(device=self.device)

    def configure(self, P, kind):
        if SOLVER_CLASS[kind] == "cg":
            P.krylov, P.preconditioner = "cg", self.cg_precond[kind]
            P.newton_solver_tolerance = self.tol
        else:
            P.krylov, P.preconditioner = "gmres", "block"
            P.gmres_restart = self.gmres_restart
            P.newton_solver_tolerance = self.ns_tol
        P.gmres_max_iters = 20000
        P.gmres_rtol = self.krylov_rtol
        P.newton_correction_tolerance = self.correction_tol
        P.max_newton_iterations = self.max_newton

    def create(self, sig, slots):
        return SlotPool(self, sig, slots)

    def solve_alone(self, request):
        """The request by itself in a pool of one slot (the alone-versus-mix check)."""
        pool = SlotPool(self, request.signature, 1, warm=False)
        pool.load(0, request)
        while not pool.step():
            pass
        return pool.take(0)
