#!/usr/bin/env python3
"""Benchmark of the assembly-and-solve path on one MI355X.

Workload (headline): the batched sweep of independent 2D Bratu problems
-lap u = lambda e^u on the unit square, 512 problems, N ~ U{16..128} per side, P1 or P2
50/50, lambda ~ U(0.5, 3.0), u = 0 on the boundary, all device-resident in one batch.
A step is one full steady Newton solve of every problem from u = 0 (pyoomph's Newton
semantics: tolerance 1e-8 on max |R|, always one step); the Jacobians are assembled by
the FP64 element kernels and the linear systems solved by batched preconditioned CG
(smoothed-aggregation multigrid V-cycle by default, Jacobi with --precond jacobi), each
problem with its own scalars and convergence test.

Reported: solves/s over the timed steps (value), p50/p99 of the per-problem completion
time inside a step, the max relative error against the FP64 CPU oracle on a sample,
FP64 assembly GB/s against the 8 TB/s HBM peak, the roofline of the dominant kernel
(HIP events on the problem's stream around every launch), the plain FP64
single-problem GPU path on a sample, and the oracle timed on the host (cpu_baseline).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Newton solves/s & p99 latency at 1e-8 rel err; assembly HBM GB/s vs peak"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nprob", type=int, default=512)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=64,
                    help="problems in the stratified CPU-baseline sample")
    ap.add_argument("--cpu-seconds", type=float, default=40.0,
                    help="bound on the CPU-baseline sample (seconds of oracle work)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--precision", choices=("f64", "mixed"), default="f64",
                    help="f64: FP64 Jacobian and Krylov; mixed: FP32 element kernels for the "
                         "Jacobian and FP32 inner CG inside FP64 iterative refinement (FP64 "
                         "residuals, updates and norms)")
    ap.add_argument("--precond", choices=("jacobi", "amg", "ilu0"), default="amg",
                    help="CG preconditioner: Jacobi, or the smoothed-aggregation multigrid V-cycle")
    ap.add_argument("--err-sample", type=int, default=6,
                    help="problems (spread over the size range, largest included) checked against "
                         "the FP64 oracle")
    ap.add_argument("--serving", default=None, metavar="RATES",
                    help="instead of the sweep: serve a seeded Poisson-arrival stream of mixed requests "
                         "at each of these comma-separated rates (req/s) and print one JSON line per rate")
    ap.add_argument("--serving-requests", type=int, default=48)
    ap.add_argument("--serving-precond", choices=("jacobi", "amg"), default="jacobi",
                    help="preconditioner of the CG class of served requests")
    ap.add_argument("--full", choices=("cavity256", "rd64", "cavity_startup"), default=None,
                    help="instead of the sweep: one full-size single-problem solve, the 256x256 "
                         "Taylor-Hood cavity (GMRES(150) + multicolour ILU(0), FP64) or the 64^3 Q2 "
                         "reaction-diffusion cube (FP32 ILU(0)-GMRES inside FP64 refinement, compared "
                         "with the FP64 CG solve of the same problem), or the start-up of the 128x128 "
                         "cavity from rest (10 BDF steps of dt 0.01)")
    ap.add_argument("--transient", type=int, default=0, metavar="STEPS",
                    help="instead of the steady sweep: STEPS BDF time steps (dt 0.02, BDF1 first) of "
                         "du/dt = lap u + lambda e^u for every problem of the sweep from u = 0")
    ap.add_argument("--continuation", type=int, default=0, metavar="STEPS",
                    help="instead of the steady sweep: STEPS pseudo-arclength steps in lambda (ds 0.5, "
                         "max 0.3) of every problem of the sweep from lambda = 0, u = 0 (GMRES + --precond)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 1 GiB read + 1 GiB write streaming kernel (three launches), the "
                         "reference for scaling HBM byte counters")
    return ap.parse_args()


def spmv_bytes(P):
    """Algorithmic bytes of one FP64 CSR SpMV over all rows: values + column indices,
    row starts, x (once per dof) and y."""
    return 12 * P.nnz + 4 * (P.ndof + 1) + 8 * P.ndof + 8 * P.ndof


def assembly_bytes(P):
    """Minimal bytes of one FP64 residual + Jacobian assembly: nodal values, coordinates,
    connectivity read once; CSR values and residual written once."""
    b = 0
    for mp in P.parts:
        m = mp.mesh
        b += 8 * m.nnode * mp.equations.nvalue + 8 * m.nnode * m.dim
        b += 4 * m.nelement * m.elements.shape[1] + 8 * mp.ndof
    return b + 8 * P.nnz


def timing(P, fam):
    import ctypes
    from pyoomph_amd import _lib
    n = ctypes.c_int64()
    ms = ctypes.c_double()
    _lib.check(_lib.load().pyhip_timing_read(P._handle, fam.encode(), ctypes.byref(n), ctypes.byref(ms)))
    return n.value, ms.value


def timing_bytes(P, fam):
    """Algorithmic bytes the library counted for the family's launches (0 if it does not
    count them)."""
    import ctypes
    from pyoomph_amd import _lib
    b = ctypes.c_int64()
    _lib.check(_lib.load().pyhip_timing_bytes(P._handle, fam.encode(), ctypes.byref(b)))
    return b.value


def copy_bandwidth():
    """Measured streaming bandwidth: a 1 GiB read + 1 GiB write elementwise kernel
    (torch.mul into a second FP64 tensor), best of 5, timed with events."""
    import torch
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.mul(a, 1.0, out=b)
    best = None
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        torch.mul(a, 1.0, out=b)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    del a, b
    return 2 * (1 << 30) / (best * 1e-3) / 1e9


def wl_ndof(spec):
    """Unknowns of a sweep problem: interior nodes of N x N P1 or P2 triangles."""
    N, sp, _ = spec
    m = 2 * N - 1 if sp == "C2" else N - 1
    return m * m


# mixed precision: Newton stops once its FP64 update is below this fraction of the
# solution as well (an FP32 Jacobian converges linearly; the update bounds the error)
MIXED_CORRECTION_TOL = 1e-10


def configure(P, precision="f64", precond="jacobi"):
    P.krylov = "cg"
    P.preconditioner = precond
    P.jacobian_precision = "f32" if precision == "mixed" else "f64"
    P.linear_precision = "f32" if precision == "mixed" else "f64"
    P.gmres_max_iters = 20000
    if precision == "mixed":
        P.newton_correction_tolerance = MIXED_CORRECTION_TOL
        P.max_newton_iterations = 20


def serving_main(args):
    """Served Newton solves: requests (2D P2 Poisson, 3D Q2 reaction-diffusion, 2D
    Taylor-Hood cavity) of 5 k - 600 k unknowns (cavity up to 12 k) arrive as a seeded
    Poisson process, each with its mesh; admission under a 200 GB HBM budget, continuous
    batching per solver class on 4 worker streams (pattern set-up, upload, Newton solve
    and download inside the latency). Per rate: solves/s over the stream (first arrival
    to last completion), p50/p99 latency (arrival to result), and the oracle's error on
    the smallest request of each kind."""
    import torch
    from pyoomph_amd.serving import GpuExecutor, ServingLoop, poisson_stream, summarize
    from oracle import newton as on
    from tests.oracle_bridge import oracle_problem
    torch.cuda.set_device(0)
    ex = GpuExecutor(cg_precond=args.serving_precond)
    # warm the library and the element kernels once (not timed)
    ServingLoop(ex, workers=1).run(poisson_stream(1000.0, 3, seed=99, dof_range=(2000, 4000),
                                                  ns_max_dofs=3000), realtime=False)
    for rate in [float(r) for r in args.serving.split(",")]:
        reqs = poisson_stream(rate, args.serving_requests, seed=args.seed,
                              dof_range=(5_000, 600_000), ns_max_dofs=12_000)
        for q in reqs:      # requests arrive with their meshes (built before the clock starts)
            q.attach_payload()
        loop = ServingLoop(ex, hbm_budget=200e9, workers=4, max_batch_dofs=4_000_000)
        torch.cuda.synchronize()
        res = loop.run(reqs, realtime=True)
        s = summarize(res)
        errs = {}
        for kind in ("poisson", "rd", "ns"):
            cand = [(q, r) for q, r in zip(sorted(reqs, key=lambda q: q.rid), res) if q.kind == kind]
            if not cand:
                continue
            q, r = min(cand, key=lambda qr: qr[0].ndof)
            mp = q.mesh_problem()
            if kind == "poisson":
                op, tol = oracle_problem("poisson", N=q.N, seed=q.seed), 1e-8
            elif kind == "rd":
                op, tol = oracle_problem("rd", N=q.N, f=mp.equations.source_field), 1e-8
            else:
                op, tol = oracle_problem("ns", N=q.N, Re=100.0), 1e-10
            xo = on.newton_solve(op, np.zeros(op.ndof), tol=tol)
            errs[kind] = float(np.max(np.abs(r.dofs - xo)) / max(np.max(np.abs(xo)), 1e-300))
        kinds = {k: sum(q.kind == k for q in reqs) for k in ("poisson", "rd", "ns")}
        out = {"metric": "served Newton solves/s & p99 latency at 1e-8 rel err", "value": s["solves_per_s"],
               "unit": "solves/s", "n_gpus": 1, "higher_is_better": True, "rate_req_per_s": rate,
               "requests": s["requests"], "p50_ms": s["p50_s"] * 1e3, "p99_ms": s["p99_s"] * 1e3,
               "mean_batch": s["mean_batch"], "batches": len(loop.batches),
               "peak_reserved_GB": loop.peak_reserved / 1e9,
               "converged": int(sum(bool(r.stats.get("converged")) for r in res)),
               "max_rel_err_vs_oracle": max(errs.values()) if errs else None, "err_by_kind": errs,
               "dtype": "f64", "data": "synthetic",
               "config": {"workload": "serving: Poisson arrivals, P2 Poisson / Q2 reaction-diffusion / "
                                      "Taylor-Hood cavity requests, 5k-600k unknowns (cavity <= 12k)",
                          "mix": kinds, "total_dofs": int(sum(q.ndof for q in reqs)), "seed": args.seed,
                          "hbm_budget_GB": 200, "workers": 4, "max_batch_dofs": 4_000_000,
                          "cg_precond": args.serving_precond}}
        print(json.dumps(out), flush=True)


def full_main(args):
    """One full-size problem, solved from zero on the device; prints one JSON line."""
    import torch
    from pyoomph_amd import _lib, workloads as wl
    from pyoomph_amd.problem import Problem
    torch.cuda.set_device(0)
    fams = ("element", "gather_values", "spmv", "precond", "vector")

    def run(P):
        P.dofs.zero_()
        _lib.check(_lib.load().pyhip_timing_enable(P._handle, 1))
        torch.cuda.synchronize()
        t = time.perf_counter()
        st = P.solve(raise_on_failure=False)[0]
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        k = {f: dict(zip(("launches", "ms"), timing(P, f))) for f in fams}
        _lib.check(_lib.load().pyhip_timing_enable(P._handle, 0))
        st = {k2: st[k2] for k2 in ("newton_iters", "linear_iters", "linear_failures", "converged",
                                    "final_maxres")}
        return el, st, k

    if args.full == "cavity_startup":
        t = time.perf_counter()
        P = Problem(wl.lid_driven_cavity(128, Re=100.0))
        setup = time.perf_counter() - t
        P.krylov, P.preconditioner, P.ilu_order, P.gmres_restart = "gmres", "ilu0", "auto", 150
        P.gmres_max_iters, P.newton_solver_tolerance = 8000, 1e-10
        P.set_initial_condition(np.zeros(P.ndof))
        _lib.check(_lib.load().pyhip_timing_enable(P._handle, 1))
        torch.cuda.synchronize()
        t = time.perf_counter()
        steps = []
        for _ in range(10):
            st = P.unsteady_newton_solve(0.01, raise_on_failure=False)[0]
            steps.append({k2: st[k2] for k2 in ("newton_iters", "linear_iters", "converged", "final_maxres")})
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        k = {f: dict(zip(("launches", "ms"), timing(P, f))) for f in fams}
        out = {"workload": "2D Taylor-Hood lid-driven cavity start-up from rest, 128x128 Q2/Q1, Re=100, "
                           "10 steps dt 0.01 (BDF1 first, then BDF2), Newton to max|R| < 1e-10 each",
               "ndof": P.ndof, "nnz": P.nnz, "setup_s": setup, "solve_s": el, "steps_per_s": 10 / el,
               "solver": "GMRES(150) + multicolour ILU(0), FP64", "steps": steps, "kernel_ms": k}
        print(json.dumps(out), flush=True)
        return
    if args.full == "cavity256":
        t = time.perf_counter()
        P = Problem(wl.lid_driven_cavity(256, Re=100.0))
        setup = time.perf_counter() - t
        P.krylov, P.preconditioner, P.ilu_order, P.gmres_restart = "gmres", "ilu0", "auto", 150
        P.gmres_max_iters, P.newton_solver_tolerance = 8000, 1e-10
        el, st, k = run(P)
        out = {"workload": "2D Taylor-Hood lid-driven cavity, 256x256 Q2/Q1, Re=100, Newton from 0 to "
                           "max|R| < 1e-10", "ndof": P.ndof, "nnz": P.nnz, "setup_s": setup,
               "solve_s": el, "solver": "GMRES(150) + multicolour ILU(0), FP64", "stats": st, "kernel_ms": k}
    else:
        t = time.perf_counter()
        P = Problem(wl.reaction_diffusion_cube(64))
        setup = time.perf_counter() - t
        # max|R| starts at 4e-7 on this mesh, so pyoomph's absolute 1e-8 stops after one
        # FP32-Jacobian step (8.5e-6 from the FP64 solution); the run goes on to 1e-15
        P.krylov, P.preconditioner = "gmres", "ilu0"
        P.jacobian_precision, P.linear_precision = "f32", "f32"
        P.gmres_max_iters, P.newton_solver_tolerance = 8000, 1e-15
        el, st, k = run(P)
        x_mixed = P.get_current_dofs()
        P.krylov, P.preconditioner = "cg", "jacobi"
        P.jacobian_precision, P.linear_precision = "f64", "f64"
        P.newton_solver_tolerance = 1e-15
        el64, st64, k64 = run(P)
        x64 = P.get_current_dofs()
        diff = float(np.max(np.abs(x_mixed - x64)) / max(np.max(np.abs(x64)), 1e-300))
        out = {"workload": "3D reaction-diffusion -lap u + u^3 = f, 64^3 Q2 bricks, f = smoothed U(-1,1) "
                           "noise, Newton from 0 to max|R| < 1e-15", "ndof": P.ndof, "nnz": P.nnz, "setup_s": setup,
               "solve_s": el, "solver": "FP32 element kernels + FP32 ILU(0)-GMRES(50) inside FP64 iterative "
                                        "refinement, FP64 residual", "stats": st, "kernel_ms": k,
               "fp64_cg_jacobi_solve_s": el64, "fp64_stats": st64,
               "max_rel_diff_vs_fp64_solve": diff}
    print(json.dumps(out), flush=True)


def transient_main(args):
    """Time stepping of the whole sweep: every step is one unsteady Newton solve of all
    512 problems (BDF2 weights, BDF1 on the first step); problem-steps/s over the timed
    steps, and the oracle's BDF Newton on three problems of the sweep over the same steps."""
    import torch
    from pyoomph_amd import workloads as wl
    from pyoomph_amd.problem import Problem
    torch.cuda.set_device(0)
    specs = wl.bratu_sweep_specs(args.nprob, seed=args.seed)
    P = Problem([wl.bratu(N, sp, lam) for N, sp, lam in specs])
    configure(P, "f64", args.precond)
    dts = [0.02] * args.transient
    P.set_initial_condition(np.zeros(P.ndof))     # warmup: plans, tuning
    P.unsteady_newton_solve(dts[0])
    P.set_initial_condition(np.zeros(P.ndof))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    newton = []
    for dt in dts:
        st = P.unsteady_newton_solve(dt)
        assert all(s["converged"] for s in st)
        newton.append(max(s["newton_iters"] for s in st))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    x = P.get_current_dofs()
    errs = []
    from oracle import timestepping as ots
    from tests.oracle_bridge import oracle_problem
    order = sorted(range(len(specs)), key=lambda i: wl_ndof(specs[i]))
    for i in (order[0], order[len(order) // 4], order[len(order) // 2]):
        N, sp, lam = specs[i]
        op = oracle_problem("bratu", N=N, space=sp, lam=lam)
        ts, xo = ots.TimeState(np.zeros(op.ndof)), np.zeros(op.ndof)
        for dt in dts:
            xo = ots.unsteady_newton_step(op, xo, ts, dt)
        xg = P.sub_dofs(i, x)
        errs.append(float(np.max(np.abs(xg - xo)) / max(np.max(np.abs(xo)), 1e-300)))
    out = {"metric": "transient Newton problem-steps/s at 1e-8 rel err", "value": args.nprob * len(dts) / elapsed,
           "unit": "problem-steps/s", "n_gpus": 1, "higher_is_better": True, "steps": len(dts),
           "ms_per_step": elapsed / len(dts) * 1e3, "newton_iters_per_step": newton,
           "max_rel_err_vs_oracle": max(errs), "dtype": "f64", "data": "synthetic",
           "config": {"workload": "bratu_sweep transient: du/dt = lap u + lambda e^u, %d problems, N~U{16..128}, "
                                  "P1/P2, dt 0.02, BDF2 (BDF1 first step), from u = 0" % args.nprob,
                      "dofs": P.ndof, "preconditioner": args.precond, "oracle_sample": "3 problems, same steps"}}
    print(json.dumps(out), flush=True)


def continuation_main(args):
    """Pseudo-arclength continuation of the whole sweep: each step is one batched
    arclength step of all problems (each with its own lambda, step and state);
    problem-steps/s, and the oracle's residual at the final point of three problems."""
    import torch
    from pyoomph_amd import workloads as wl
    from pyoomph_amd.problem import Problem
    from oracle.newton import max_abs
    from tests.oracle_bridge import oracle_problem
    torch.cuda.set_device(0)
    specs = wl.bratu_sweep_specs(args.nprob, seed=args.seed)
    P = Problem([wl.bratu(N, sp, 0.0) for N, sp, _ in specs])
    P.krylov, P.preconditioner, P.gmres_restart = "gmres", args.precond, 100
    ds = np.full(len(specs), 0.5)
    P.arclength_continuation("lambda", ds, max_ds=0.3)     # warmup: plans, tuning
    P.set_current_dofs(np.zeros(P.ndof))
    P.reset_arc_length_parameters()
    for i in range(len(specs)):
        P.set_global_parameter(i, 0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = []
    for _ in range(args.continuation):
        ds = P.arclength_continuation("lambda", ds, max_ds=0.3)
        its.append(max(a["newton_iters"] for a in P.arclength_stats()))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    lam = P.global_parameters()
    x = P.get_current_dofs()
    order = sorted(range(len(specs)), key=lambda i: wl_ndof(specs[i]))
    res = []
    for i in (order[0], order[len(order) // 4], order[len(order) // 2]):
        N, sp, _ = specs[i]
        op = oracle_problem("bratu", N=N, space=sp, lam=float(lam[i]))
        res.append(max_abs(op.get_residuals(P.sub_dofs(i, x))))
    out = {"metric": "arclength continuation problem-steps/s", "value": args.nprob * args.continuation / elapsed,
           "unit": "problem-steps/s", "n_gpus": 1, "higher_is_better": True, "steps": args.continuation,
           "ms_per_step": elapsed / args.continuation * 1e3, "max_newton_iters_per_step": its,
           "lambda_range_after": [float(lam.min()), float(lam.max())],
           "oracle_max_residual_at_final_points": max(res), "dtype": "f64", "data": "synthetic",
           "config": {"workload": "bratu_sweep continuation: %d problems, N~U{16..128}, P1/P2, from lambda = 0, "
                                  "ds 0.5 (max 0.3), GMRES(100) + %s" % (args.nprob, args.precond), "dofs": P.ndof}}
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.continuation:
        continuation_main(args)
        return
    if args.transient:
        transient_main(args)
        return
    if args.full:
        full_main(args)
        return
    if args.serving:
        if args.gpus != 1:
            print("bench.py measures one GPU: --gpus %d is not supported" % args.gpus, file=sys.stderr)
            sys.exit(2)
        serving_main(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from pyoomph_amd import _lib, workloads as wl
    from pyoomph_amd.problem import Problem

    torch.cuda.set_device(0)
    if args.copy_calibration:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        del a, b
    specs = wl.bratu_sweep_specs(args.nprob, seed=args.seed)
    parts = [wl.bratu(N, sp, lam) for N, sp, lam in specs]
    P = Problem(parts)
    configure(P, args.precision, args.precond)
    print("# batch: %d problems, %d dofs, %d nnz" % (len(parts), P.ndof, P.nnz), file=sys.stderr, flush=True)

    def step():
        P.dofs.zero_()
        return P.solve()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    _lib.check(_lib.load().pyhip_timing_enable(P._handle, 1))
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = step()
        lat += [s["done_ms"] for s in st]
        assert all(s["converged"] for s in st)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    fams = {f: timing(P, f) for f in ("element", "gather_values", "gather_residual", "spmv",
                                       "precond", "vector")}
    spmv_work = timing_bytes(P, "spmv")
    _lib.check(_lib.load().pyhip_timing_enable(P._handle, 0))
    x_gpu = P.get_current_dofs()
    solves_per_s = args.nprob * args.steps / elapsed
    ms_per_step = elapsed / args.steps * 1e3

    # FP64 assembly (element kernels + gathers), timed with the same events
    _lib.check(_lib.load().pyhip_timing_enable(P._handle, 1))
    for _ in range(5):
        P.assemble_device("f64")
    torch.cuda.synchronize()
    asm_ms = sum(timing(P, f)[1] for f in ("element", "gather_values", "gather_residual")) / 5
    asm_el = timing(P, "element")[1] / 5
    _lib.check(_lib.load().pyhip_timing_enable(P._handle, 0))
    asm_gbs = assembly_bytes(P) / (asm_ms * 1e-3) / 1e9
    copy_gbs = copy_bandwidth()

    # dominant kernel family of the timed region and its roofline ("precond" times whole
    # preconditioner applies, a composite whose spmv launches are in the "spmv" family)
    dom = max((f for f in fams if f != "precond"), key=lambda f: fams[f][1])
    n_l, ms_l = fams[dom]
    avg_ms = ms_l / max(n_l, 1)
    if dom == "spmv":
        # the CG visits only the rows of problems still iterating, and the multigrid levels
        # are smaller than the matrix: the library counts the bytes each launch processed
        alg = spmv_work / max(n_l, 1) if spmv_work else spmv_bytes(P)
    elif dom in ("element", "gather_values", "gather_residual"):
        alg = assembly_bytes(P)
    else:
        alg = 16 * P.ndof   # two vectors read per launch (dot products / updates)
    achieved = alg / (avg_ms * 1e-3) / 1e9
    traffic = None
    # measured HBM traffic of this kernel family on this configuration (two rocprofv3 PMC passes)
    tfile = os.path.join(ROOT, "profiles", "traffic_%s_bratu_sweep%s.txt"
                         % (dom, "" if args.precond == "amg" else "_" + args.precond))
    if os.path.exists(tfile) and args.precision == "f64":
        for line in open(tfile):
            if line.startswith("traffic_bytes_per_launch="):
                traffic = float(line.split("=", 1)[1])

    # plain FP64 single-problem GPU path on a sample (one problem at a time, CG + Jacobi),
    # and the same sample one problem at a time with the batch's own solver, so that the
    # batching gain and the preconditioner gain are reported apart
    single = single_same = None
    if not args.no_single:
        def one_at_a_time(precision, precond):
            t_single, n_single = 0.0, 0
            for N, sp, lam in specs[:8]:
                Q = Problem(wl.bratu(N, sp, lam))
                configure(Q, precision, precond)
                Q.solve()                      # pattern / plan set-up outside the clock
                Q.dofs.zero_()
                torch.cuda.synchronize()
                t = time.perf_counter()
                Q.solve()
                torch.cuda.synchronize()
                t_single += time.perf_counter() - t
                n_single += 1
                del Q
            return n_single / t_single
        single = one_at_a_time("f64", "jacobi")
        single_same = one_at_a_time(args.precision, args.precond)

    # CPU-baseline leg: the oracle checks accuracy (problems spread over the size range,
    # largest included, plus the timed sample) and is timed on the smallest problems
    cpu = None
    errs = []
    if not args.no_cpu_baseline:
        from oracle import newton as on
        from tests.oracle_bridge import oracle_problem
        order = sorted(range(len(specs)), key=lambda i: (specs[i][0] * (2 if specs[i][1] == "C2" else 1)))

        def err_of(i, xo):
            xg = P.sub_dofs(i, x_gpu)
            return float(np.max(np.abs(xg - xo)) / max(np.max(np.abs(xo)), 1e-300))

        if args.err_sample > 0:
            picks = sorted(set(order[int(round(k * (len(order) - 1) / max(args.err_sample - 1, 1)))]
                               for k in range(args.err_sample)))
            for i in picks:
                N, sp, lam = specs[i]
                op = oracle_problem("bratu", N=N, space=sp, lam=lam)
                errs.append(err_of(i, on.newton_solve(op, np.zeros(op.ndof))))
            print("# accuracy sample: %s" % [(specs[i][0], specs[i][1]) for i in picks], file=sys.stderr,
                  flush=True)
        # timed sample: a stratified slice of the sweep (every len/k-th problem in order of
        # unknowns, k = --cpu-sample), so its solves/s is that of the sweep's size mix
        k = min(args.cpu_sample, len(order))
        strat = [order[int((j + 0.5) * len(order) / k)] for j in range(k)]
        n_cpu, t_cpu = 0, 0.0
        for i in strat:
            N, sp, lam = specs[i]
            op = oracle_problem("bratu", N=N, space=sp, lam=lam)
            t = time.perf_counter()
            xo = on.newton_solve(op, np.zeros(op.ndof))
            t_cpu += time.perf_counter() - t
            n_cpu += 1
            errs.append(err_of(i, xo))
            if t_cpu > args.cpu_seconds:
                break
        done = strat[:n_cpu]
        cpu = {"value": n_cpu / t_cpu, "unit": "solves/s", "cores": 1, "kind": "port",
               "host_cores": os.cpu_count(),
               "sample": "%d of %d problems of a stratified slice of the sweep (every %d-th by unknowns, "
                         "%d..%d unknowns), oracle C element code + vectors_of_pairs CSR + scipy SuperLU, "
                         "one thread, %.1f s" % (n_cpu, k, len(order) // k,
                                                 min(wl_ndof(specs[i]) for i in done),
                                                 max(wl_ndof(specs[i]) for i in done), t_cpu)}

    max_err = max(errs) if errs else None
    lat = np.array(lat)
    out = {
        "metric": METRIC,
        "value": solves_per_s,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "f64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "bratu_sweep: %d independent -lap u = lambda e^u problems, "
                               "N~U{16..128}, P1/P2 50/50, lambda~U(0.5,3)" % args.nprob,
                   "problems": args.nprob, "dofs": P.ndof, "nnz": P.nnz,
                   "solver": ("Newton + batched CG/%s, FP64" % args.precond if args.precision == "f64" else
                              "Newton (FP64 residual, stops at max|R| < 1e-8 and max|dx| <= %g max|x|) + "
                              "FP32 element kernels for J, FP32 batched CG/%s inside FP64 iterative refinement"
                              % (MIXED_CORRECTION_TOL, args.precond)),
                   "precision": args.precision, "preconditioner": args.precond, "seed": args.seed},
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "max_rel_err_vs_oracle": max_err,
        "assembly_fp64_ms": asm_ms,
        "assembly_element_kernel_ms": asm_el,
        "assembly_GBps": asm_gbs,
        "assembly_frac_of_hbm_peak": asm_gbs / HBM_PEAK_GBS,
        "copy_kernel_GBps": copy_gbs,
        "assembly_frac_of_copy": asm_gbs / copy_gbs,
        "single_problem_fp64_solves_per_s": single,
        "batched_speedup_vs_single": (solves_per_s / single) if single else None,
        "single_problem_same_solver_solves_per_s": single_same,
        "batched_speedup_vs_single_same_solver": (solves_per_s / single_same) if single_same else None,
        "kernel_ms": {f: {"launches": v[0], "ms": v[1]} for f, v in fams.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": alg, "avg_launch_ms": avg_ms,
                     "share_of_step": ms_l / (elapsed * 1e3), "frac_of_copy": achieved / copy_gbs,
                     "full_matrix_bytes": spmv_bytes(P)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
