#!/usr/bin/env python3
"""Benchmark of the assembly-and-solve path on one MI355X.

Workload (headline): the batched sweep of independent 2D Bratu problems
-lap u = lambda e^u on the unit square, 512 problems, N ~ U{16..128} per side, P1 or P2
50/50, lambda ~ U(0.5, 3.0), u = 0 on the boundary, all device-resident in one batch.
A step is one full steady Newton solve of every problem from u = 0: pyoomph's Newton
(max |R| < 1e-8, always one step) with, in addition, an FP64 update below 1e-6 of the
solution (on fine meshes a small residual alone does not make a small error), inexact
linear solves (batched preconditioned CG to 5e-3 relative residual per Newton step:
smoothed-aggregation multigrid V-cycle by default, Jacobi with --precond jacobi), each
problem with its own scalars and convergence tests. FP64 (the value), and mixed precision
beside it: FP32 element kernels for the Jacobian, one FP32 CG solve with the FP32 V-cycle
per Newton step, FP64 residual and norms (the Newton loop is the FP64 refinement).

Reported: solves/s over the timed steps (value), p50/p99 of the per-problem completion
time inside a step, the max relative error against the converged FP64 CPU oracle
(SuperLU Newton to max|R| < 1e-13 and one step more) on problems across the size range, FP64
assembly GB/s
against the 8 TB/s HBM peak, the roofline of the dominant kernel (HIP events on the
problem's stream around every launch), the single-problem GPU paths on a sample (plain
FP64 CG + Jacobi, and the batch's own solver), and the reference's CPU path as a per-core
process farm on the host (cpu_baseline).
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
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="bound on the CPU-baseline sample (seconds of oracle work)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-other", action="store_true", help="skip the run in the other precision")
    ap.add_argument("--precision", choices=("f64", "mixed"), default="f64",
                    help="f64: FP64 Jacobian and Krylov; mixed: FP32 element kernels for the "
                         "Jacobian and FP32 inner CG inside FP64 iterative refinement (FP64 "
                         "residuals, updates and norms)")
    ap.add_argument("--precond", choices=("jacobi", "amg", "ilu0"), default="amg",
                    help="CG preconditioner: Jacobi, or the smoothed-aggregation multigrid V-cycle")
    ap.add_argument("--krylov-rtol", type=float, default=None,
                    help="relative tolerance of each linear solve of the sweep (default 5e-3)")
    ap.add_argument("--correction-tol", type=float, default=None,
                    help="Newton also stops only once max|dx| <= this * max|x| (default: per precision)")
    ap.add_argument("--err-sample", type=int, default=6,
                    help="problems (spread over the size range, largest included) checked against "
                         "the FP64 oracle")
    ap.add_argument("--serving", default=None, metavar="RATES",
                    help="instead of the sweep: serve a seeded Poisson-arrival stream of mixed requests "
                         "at each of these comma-separated rates (req/s) and print one JSON line per rate")
    ap.add_argument("--serving-requests", type=int, default=500)
    ap.add_argument("--serving-rtol", type=float, default=1e-4,
                    help="relative tolerance of each served linear solve")
    ap.add_argument("--serving-ctol", type=float, default=1e-7,
                    help="served Newton solves stop once max|dx| <= this * max|x| (and max|R| < tol)")
    ap.add_argument("--serving-precond", choices=("auto", "jacobi", "amg"), default="auto",
                    help="preconditioner of the CG class of served requests")
    ap.add_argument("--full", choices=("cavity256", "rd64", "cavity_startup"), default=None,
                    help="instead of the sweep: one full-size single-problem solve, the 256x256 "
                         "Taylor-Hood cavity (GMRES(100) + the block preconditioner, FP64) or the 64^3 Q2 "
                         "reaction-diffusion cube (FP32 ILU(0)-GMRES inside FP64 refinement, compared "
                         "with the FP64 CG solve of the same problem), or the start-up of the 128x128 "
                         "cavity from rest (10 BDF steps of dt 0.01)")
    ap.add_argument("--transient", type=int, default=0, metavar="STEPS",
                    help="instead of the steady sweep: STEPS BDF time steps (dt 0.02, BDF1 first) of "
                         "du/dt = lap u + lambda e^u for every problem of the sweep from u = 0")
    ap.add_argument("--continuation", type=int, default=0, metavar="STEPS",
                    help="instead of the steady sweep: STEPS pseudo-arclength steps in lambda (ds 0.5, "
                         "max 0.3) of every problem of the sweep from lambda = 0, u = 0 (GMRES + --precond)")
    ap.add_argument("--assembly-study", default=None, metavar="LIB",
                    help="compare the FP64 assembly of this build with LIB, another build of the library "
                         "(e.g. make -C pyoomph_amd/csrc EXACT=-ffp-contract=fast OUT=...): element-kernel "
                         "time and the Jacobian's zero-drop decisions and values on the Poisson 64x64 P2 "
                         "problem and the Bratu sweep")
    ap.add_argument("--assembly-study-worker", default=None, metavar="DIR", help=argparse.SUPPRESS)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 1 GiB read + 1 GiB write streaming kernel (three launches), the "
                         "reference for scaling HBM byte counters")
    return ap.parse_args()


def spmv_bytes(P):
    """Algorithmic bytes of one FP64 CSR SpMV over all rows with int32 column indices:
    values + column indices, row starts, x (once per dof) and y (the library counts its own
    launches, with 2-byte column offsets where it uses them)."""
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


# sources that decide the spmv family's launches and their bytes on the sweep
LAUNCH_SOURCES = ("spmv.h", "krylov.hip", "amg.hip", "amg_plan.cpp", "amg_plan.h")


def launch_sources_sha256():
    """SHA-256 over LAUNCH_SOURCES (in pyoomph_amd/csrc, in that order)."""
    import hashlib
    h = hashlib.sha256()
    for f in LAUNCH_SOURCES:
        h.update(open(os.path.join(ROOT, "pyoomph_amd", "csrc", f), "rb").read())
    return h.hexdigest()


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


# Newton stopping: pyoomph's max|R| < 1e-8 and, since a small residual does not make a small
# error on fine meshes (a smooth error e leaves a residual ~h^2 e), an FP64 update below
# CORRECTION_TOL of the solution. Inexact Newton: each linear solve stops at KRYLOV_RTOL
# relative residual (FP64 CG, or one FP32 CG solve whose FP64 refinement is the Newton loop).
CORRECTION_TOL = 1e-6
KRYLOV_RTOL = 5e-3
# accuracy reference: the oracle's Newton run to max|R| < 1e-13 and one step more (converged)
ORACLE_TOL = 1e-13


def configure(P, precision="f64", precond="jacobi", correction_tol=CORRECTION_TOL, inexact=True):
    """Sweep solver settings. inexact=False keeps pyoomph's Newton (max|R| < 1e-8 alone)
    with tight linear solves: the transient and continuation modes compare step by step
    with the oracle's Newton iterates."""
    P.krylov = "cg"
    P.preconditioner = precond
    P.jacobian_precision = "f32" if precision == "mixed" else "f64"
    P.linear_precision = "f32" if precision == "mixed" else "f64"
    P.gmres_max_iters = 20000
    if inexact:
        P.gmres_rtol = KRYLOV_RTOL
        P.rtol_f32 = KRYLOV_RTOL
        P.max_refinements = 1
        P.newton_correction_tolerance = correction_tol
        P.max_newton_iterations = 20


def serving_main(args):
    """Served Newton solves: requests (2D P2 Poisson, 3D Q2 reaction-diffusion, 2D
    Taylor-Hood cavity) on the served mesh sizes (5 k - 490 k unknowns, cavity 6 k - 11 k)
    arrive as a seeded Poisson process, each with its own data. The server starts with one
    slot pool per mesh class (mesh, pattern, solver plans and workspace on the device,
    within a 200 GB HBM budget; start-up reported, not timed); requests join their pool at
    Newton-step boundaries and leave when converged (slot load, Newton steps and download
    inside the latency), at most 4 pools stepping at once. Per rate: solves/s over the
    stream (first arrival to last completion), p50/p99 latency (arrival to result), mean
    slot occupancy, and the error against the oracle's converged solution on a sample of
    requests (the smallest and one mid-size request of each kind)."""
    import torch
    from pyoomph_amd.serving import (SIZES, GpuPools, ServingLoop, attach_payloads, ns_reynolds, poisson_stream,
                                     summarize)
    from oracle import newton as on
    from tests.oracle_bridge import oracle_problem
    torch.cuda.set_device(0)
    ex = GpuPools(cg_precond=None if args.serving_precond == "auto" else args.serving_precond,
                  krylov_rtol=args.serving_rtol, correction_tol=args.serving_ctol)
    loop = ServingLoop(ex, hbm_budget=200e9, workers=4)
    t0 = time.perf_counter()
    sigs = [(k, N) for k, v in SIZES.items() for N in v]
    for sig in sigs:
        loop.prepare([sig])
        print("pool %s-%d ready (%d slots, %.2f GB)" % (sig[0], sig[1], loop.pool_of[sig].slots,
                                                       loop.pool_of[sig].device_bytes / 1e9), file=sys.stderr, flush=True)
    startup = time.perf_counter() - t0
    for rate in [float(r) for r in args.serving.split(",")]:
        reqs = attach_payloads(poisson_stream(rate, args.serving_requests, seed=args.seed))
        loop.steps = {sig: [] for sig in loop.pool_of}
        torch.cuda.synchronize()
        res = loop.run(reqs, realtime=True)
        s = summarize(res)
        errs = {}
        byrid = sorted(reqs, key=lambda q: q.rid)
        for kind in ("poisson", "rd", "ns"):
            cand = sorted([(q.ndof, q.rid) for q in byrid if q.kind == kind])
            pick = {cand[0][1]} if cand else set()
            mid = [rid for nd, rid in cand if nd <= 35_000]
            if mid:
                pick.add(mid[-1])
            for rid in sorted(pick):
                q, r = byrid[rid], res[rid]
                if kind == "poisson":
                    op = oracle_problem("poisson", N=q.N, seed=q.seed)
                elif kind == "rd":
                    op = oracle_problem("rd", N=q.N, f=q.payload[1])
                else:
                    op = oracle_problem("ns", N=q.N, Re=ns_reynolds(q.seed))
                xo = on.converged_solve(op, np.zeros(op.ndof), tol=1e-12)
                e = float(np.max(np.abs(r.dofs - xo)) / max(np.max(np.abs(xo)), 1e-300))
                errs["%s-%d" % (kind, q.N)] = max(e, errs.get("%s-%d" % (kind, q.N), 0.0))
        kinds = {k: sum(q.kind == k for q in reqs) for k in ("poisson", "rd", "ns")}
        occ = loop.occupancy()
        busy = sum(v["mean_busy"] * v["steps"] for v in occ.values())
        steps = sum(v["steps"] for v in occ.values())
        out = {"metric": "served Newton solves/s & p99 latency at 1e-8 rel err", "value": s["solves_per_s"],
               "unit": "solves/s", "n_gpus": 1, "higher_is_better": True, "rate_req_per_s": rate,
               "requests": s["requests"], "p50_ms": s["p50_s"] * 1e3, "p99_ms": s["p99_s"] * 1e3,
               "mean_newton_steps": s["mean_newton_steps"], "converged": s["converged"],
               "pool_steps": steps, "mean_busy_slots_per_step": busy / max(steps, 1), "pools": occ,
               "max_rel_err_vs_oracle": max(errs.values()) if errs else None, "err_sample": errs,
               "dtype": "f64", "data": "synthetic",
               "config": {"workload": "serving: Poisson arrivals, P2 Poisson / Q2 reaction-diffusion / "
                                      "Taylor-Hood cavity requests on 16 mesh classes, 5k-490k unknowns",
                          "mix": kinds, "total_dofs": int(sum(q.ndof for q in reqs)), "seed": args.seed,
                          "hbm_budget_GB": 200, "pools": len(loop.pool_of),
                          "pool_GB": round(loop.reserved / 1e9, 2), "startup_s": round(startup, 1),
                          "workers": loop.workers, "cg_precond": ex.cg_precond,
                          "solver": "inexact Newton (Krylov rtol %g, stop at max|R| < 1e-8 (cavity 1e-10) and "
                                    "max|dx| <= %g max|x|)" % (ex.krylov_rtol, ex.correction_tol)}}
        print(json.dumps(out), flush=True)
    loop.close()


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
        # ILU(0): with the BDF mass term the Schur complement is no longer the pressure mass
        # over the viscosity, so the block preconditioner is for steady solves
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
        P.krylov, P.preconditioner, P.gmres_restart = "gmres", "block", 100
        P.gmres_max_iters, P.newton_solver_tolerance = 8000, 1e-10
        el, st, k = run(P)
        out = {"workload": "2D Taylor-Hood lid-driven cavity, 256x256 Q2/Q1, Re=100, Newton from 0 to "
                           "max|R| < 1e-10", "ndof": P.ndof, "nnz": P.nnz, "setup_s": setup,
               "solve_s": el, "solver": "GMRES(100) + block-triangular preconditioner (multigrid V-cycle on the "
                                        "velocity components, lumped pressure mass / viscosity for the Schur "
                                        "complement), FP64", "stats": st, "kernel_ms": k}
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
        # the same mixed precision with CG + Jacobi as the FP32 inner solver
        P.krylov, P.preconditioner = "cg", "jacobi"
        P.jacobian_precision, P.linear_precision = "f32", "f32"
        P.newton_solver_tolerance = 1e-15
        el_cg, st_cg, _ = run(P)
        x_cg = P.get_current_dofs()
        diff_cg = float(np.max(np.abs(x_cg - x64)) / max(np.max(np.abs(x64)), 1e-300))
        # FP64 CG with the multigrid (its element level: Q1 functions on the brick vertices)
        P.krylov, P.preconditioner = "cg", "amg"
        P.jacobian_precision, P.linear_precision = "f64", "f64"
        P.newton_solver_tolerance = 1e-15
        el_amg, st_amg, _ = run(P)
        diff_amg = float(np.max(np.abs(P.get_current_dofs() - x64)) / max(np.max(np.abs(x64)), 1e-300))
        out = {"workload": "3D reaction-diffusion -lap u + u^3 = f, 64^3 Q2 bricks, f = smoothed U(-1,1) "
                           "noise, Newton from 0 to max|R| < 1e-15", "ndof": P.ndof, "nnz": P.nnz, "setup_s": setup,
               "solve_s": el, "solver": "FP32 element kernels + FP32 ILU(0)-GMRES(50) inside FP64 iterative "
                                        "refinement, FP64 residual", "stats": st, "kernel_ms": k,
               "fp64_cg_jacobi_solve_s": el64, "fp64_stats": st64,
               "max_rel_diff_vs_fp64_solve": diff,
               "mixed_cg_jacobi_solve_s": el_cg, "mixed_cg_jacobi_stats": st_cg,
               "mixed_cg_jacobi_max_rel_diff_vs_fp64_solve": diff_cg,
               "fp64_cg_amg_solve_s": el_amg, "fp64_cg_amg_stats": st_amg,
               "fp64_cg_amg_max_rel_diff_vs_fp64_cg_jacobi": diff_amg}
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
    configure(P, "f64", args.precond, inexact=False)
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


def assembly_study_worker(args):
    """One build's side of the assembly study, in a process of its own (the library comes
    from PYOOMPH_HIP_LIB when set). Per case: the element kernels' time per FP64 Jacobian
    assembly (mean of 5) and the Jacobian in the multi-assembly CSR (columns ascending,
    exact-zero element contributions dropped). The first build saves its kept-entry mask and
    values under DIR; a later build compares its own with them."""
    import torch
    from pyoomph_amd import _lib, workloads as wl
    from pyoomph_amd.problem import Problem
    torch.cuda.set_device(0)
    out = args.assembly_study_worker
    ref = os.path.exists(os.path.join(out, "poisson64_mask.npy"))
    cases = [("poisson64", lambda: [wl.poisson_gaussians(64, 0)]),
             ("bratu_sweep", lambda: [wl.bratu(N, sp, lam) for N, sp, lam in wl.bratu_sweep_specs(512, seed=0)])]
    L = _lib.load()
    for name, parts in cases:
        P = Problem(parts())
        P.set_current_dofs(np.random.default_rng(1).uniform(0.0, 0.5, P.ndof))
        P.assemble_device("f64")
        _lib.check(L.pyhip_timing_enable(P._handle, 1))
        for _ in range(5):
            P.assemble_device("f64")
        el = timing(P, "element")
        _lib.check(L.pyhip_timing_enable(P._handle, 0))
        J, = P.start_multiassembly().J().assemble()
        rs, col = P.structural_pattern()
        rows = np.repeat(np.arange(P.ndof, dtype=np.int64), np.diff(rs))
        jrows = np.repeat(np.arange(P.ndof, dtype=np.int64), np.diff(J.indptr))
        mask = np.isin(rows * P.ndof + col, jrows * P.ndof + J.indices)
        vals = np.zeros(P.nnz)
        vals[mask] = J.data
        r = {"case": name, "lib": _lib.LIB_PATH, "ndof": P.ndof, "structural_nnz": int(P.nnz),
             "kept_nnz": int(J.nnz), "element_launches_per_assembly": el[0] / 5, "element_ms_per_assembly": el[1] / 5}
        if not ref:
            np.save(os.path.join(out, name + "_mask.npy"), np.packbits(mask))
            np.save(os.path.join(out, name + "_vals.npy"), vals)
        else:
            m0 = np.unpackbits(np.load(os.path.join(out, name + "_mask.npy")))[:P.nnz].astype(bool)
            v0 = np.load(os.path.join(out, name + "_vals.npy"), mmap_mode="r")
            both = m0 & mask
            d = np.abs(vals[both] - v0[both])
            r.update({"decisions_changed": int(np.count_nonzero(m0 != mask)),
                      "kept_here_only": int(np.count_nonzero(mask & ~m0)),
                      "kept_there_only": int(np.count_nonzero(m0 & ~mask)),
                      "values_differing": int(np.count_nonzero(d)),
                      "max_rel_diff": float(d.max() / max(np.abs(v0[both]).max(), 1e-300)) if d.size else 0.0})
        print(json.dumps(r), flush=True)
        del P


def assembly_study(args):
    """Runs the worker on the in-tree build, then on args.assembly_study, and prints one
    JSON line: element-kernel time of each and what changed in the Jacobian."""
    import subprocess
    import tempfile
    lines = []
    with tempfile.TemporaryDirectory() as d:
        for lib in (None, args.assembly_study):
            env = dict(os.environ)
            env.pop("PYOOMPH_HIP_LIB", None)
            if lib:
                env["PYOOMPH_HIP_LIB"] = os.path.abspath(lib)
            print("# assembly study: %s" % (lib or "in-tree build"), file=sys.stderr, flush=True)
            r = subprocess.run([sys.executable, os.path.abspath(__file__), "--assembly-study-worker", d],
                               env=env, stdout=subprocess.PIPE, text=True)
            if r.returncode:
                sys.exit(r.returncode)
            lines.append([json.loads(x) for x in r.stdout.splitlines() if x.startswith("{")])
    out = {"study": "FP64 assembly: in-tree build vs %s" % args.assembly_study, "cases": []}
    for a, b in zip(*lines):
        out["cases"].append({"case": a["case"], "ndof": a["ndof"], "structural_nnz": a["structural_nnz"],
                             "kept_nnz": [a["kept_nnz"], b["kept_nnz"]],
                             "element_ms_per_assembly": [a["element_ms_per_assembly"], b["element_ms_per_assembly"]],
                             "decisions_changed": b["decisions_changed"], "kept_by_other_only": b["kept_here_only"],
                             "kept_by_in_tree_only": b["kept_there_only"], "values_differing": b["values_differing"],
                             "max_rel_diff": b["max_rel_diff"]})
    print(json.dumps(out), flush=True)


def main():
    global KRYLOV_RTOL
    args = parse()
    if args.assembly_study_worker:
        assembly_study_worker(args)
        return
    if args.assembly_study:
        assembly_study(args)
        return
    if args.krylov_rtol is not None:
        KRYLOV_RTOL = args.krylov_rtol
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
    ctol = CORRECTION_TOL if args.correction_tol is None else args.correction_tol
    print("# batch: %d problems, %d dofs, %d nnz" % (len(parts), P.ndof, P.nnz), file=sys.stderr, flush=True)
    order = sorted(range(len(specs)), key=lambda i: wl_ndof(specs[i]))

    def timed_sweep(precision, with_timing):
        """warmup + timed steps of the whole sweep in one precision; returns
        (solves/s, ms/step, latencies, stats, kernel families, spmv bytes, dofs). With
        with_timing, the timed steps are repeated once more with HIP events on the problem's
        stream around every kernel family (the per-launch durations of the roofline; the
        events cost ~8 % of a step, so the solves/s come from the first, event-free pass)."""
        configure(P, precision, args.precond, correction_tol=ctol)

        def step():
            P.dofs.zero_()
            return P.solve()

        def timed():
            lat = []
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                st = step()
                lat += [s["done_ms"] for s in st]
                assert all(s["converged"] for s in st)
            torch.cuda.synchronize()
            return time.perf_counter() - t0, lat, st

        for _ in range(args.warmup):
            step()
        elapsed, lat, st = timed()
        fams, work = None, None
        if with_timing:
            _lib.check(_lib.load().pyhip_timing_enable(P._handle, 1))
            timed()
            fams = {f: timing(P, f) for f in ("element", "gather_values", "gather_residual", "spmv",
                                               "precond", "vector")}
            work = timing_bytes(P, "spmv")
            _lib.check(_lib.load().pyhip_timing_enable(P._handle, 0))
        return (args.nprob * args.steps / elapsed, elapsed / args.steps * 1e3, np.array(lat), st, fams, work,
                P.get_current_dofs())

    # the headline precision, timed with the kernel-family events; then the other precision
    sps, ms_per_step, lat, st, fams, spmv_work, x_gpu = timed_sweep(args.precision, True)
    other = "mixed" if args.precision == "f64" else "f64"
    other_run = None if args.no_other else timed_sweep(other, False)
    configure(P, args.precision, args.precond, correction_tol=ctol)

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
        # are smaller than the matrix: the library counts the bytes the launches processed
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
    traffic_note = None
    if os.path.exists(tfile) and args.precision == "f64":
        # the profile names the sources it was measured with (the spmv kernels and the
        # solver and multigrid code that decide the launches); a figure taken with other
        # sources is not reported
        src = launch_sources_sha256()
        rec, tsrc = None, None
        for line in open(tfile):
            if line.startswith("traffic_bytes_per_launch="):
                rec = float(line.split("=", 1)[1])
            if line.startswith("launch_sources_sha256="):
                tsrc = line.split("=", 1)[1].strip()
        if rec is not None and tsrc == src:
            traffic = rec
        elif rec is not None:
            traffic_note = "the traffic profile was measured with other spmv or solver sources; not reported"

    # single-problem GPU paths on a sample (one problem per Problem, pattern and plan set up
    # outside the clock, the same stopping rules): the plain FP64 path (CG + Jacobi) and the
    # batch's own solver -- the batching gain and the preconditioner gain apart
    single = single_same = None
    if not args.no_single:
        def one_at_a_time(precision, precond):
            t_single, n_single = 0.0, 0
            for N, sp, lam in specs[:8]:
                Q = Problem(wl.bratu(N, sp, lam))
                configure(Q, precision, precond, correction_tol=ctol)
                Q.solve()
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

    # accuracy: every sampled problem against the converged FP64 oracle (SuperLU Newton to
    # max|R| < ORACLE_TOL and one step more), problems spread over the size range, the
    # largest included
    errs, errs_other = [], []
    if args.err_sample > 0:
        from oracle import newton as on
        from tests.oracle_bridge import oracle_problem
        picks = sorted(set(order[int(round(k * (len(order) - 1) / max(args.err_sample - 1, 1)))]
                           for k in range(args.err_sample)))
        for i in picks:
            N, sp, lam = specs[i]
            op = oracle_problem("bratu", N=N, space=sp, lam=lam)
            xo = on.converged_solve(op, np.zeros(op.ndof), tol=ORACLE_TOL)
            nrm = max(np.max(np.abs(xo)), 1e-300)
            errs.append(float(np.max(np.abs(P.sub_dofs(i, x_gpu) - xo)) / nrm))
            if other_run is not None:
                errs_other.append(float(np.max(np.abs(P.sub_dofs(i, other_run[6]) - xo)) / nrm))
        print("# accuracy sample: %s" % [(specs[i][0], specs[i][1]) for i in picks], file=sys.stderr, flush=True)

    cpu = None if args.no_cpu_baseline else cpu_farm_baseline(specs, order, args)

    out = {
        "metric": METRIC,
        "value": sps,
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
                   "solver": solver_text(args.precision, args.precond, ctol),
                   "precision": args.precision, "preconditioner": args.precond, "seed": args.seed},
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "max_rel_err_vs_oracle": max(errs) if errs else None,
        "oracle": "SuperLU Newton (the oracle) to max|R| < %g plus one step, on %d problems across the size "
                  "range" % (ORACLE_TOL, len(errs)),
        "assembly_fp64_ms": asm_ms,
        "assembly_element_kernel_ms": asm_el,
        "assembly_GBps": asm_gbs,
        "assembly_frac_of_hbm_peak": asm_gbs / HBM_PEAK_GBS,
        "copy_kernel_GBps": copy_gbs,
        "assembly_frac_of_copy": asm_gbs / copy_gbs,
        "single_problem_fp64_solves_per_s": single,
        "batched_speedup_vs_single": (sps / single) if single else None,
        "single_problem_same_solver_solves_per_s": single_same,
        "batched_speedup_vs_single_same_solver": (sps / single_same) if single_same else None,
        "kernel_ms": {f: {"launches": v[0], "ms": v[1]} for f, v in fams.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": os.path.relpath(tfile, ROOT) if traffic is not None else traffic_note,
                     "algorithmic_bytes_per_launch": alg, "avg_launch_ms": avg_ms,
                     "share_of_step": ms_l / (ms_per_step * args.steps),
                     "frac_of_copy": achieved / copy_gbs, "full_matrix_bytes": spmv_bytes(P)},
        "cpu_baseline": cpu,
    }
    if other_run is not None:
        o_sps, o_ms, o_lat = other_run[0], other_run[1], other_run[2]
        out[other + "_precision" if other == "mixed" else "fp64_precision"] = {
            "value": o_sps, "ms_per_step": o_ms, "p50_ms": float(np.percentile(o_lat, 50)),
            "p99_ms": float(np.percentile(o_lat, 99)),
            "max_rel_err_vs_oracle": max(errs_other) if errs_other else None,
            "speedup_vs_single_problem_fp64": (o_sps / single) if single else None,
            "solver": solver_text(other, args.precond, ctol)}
    print(json.dumps(out))


def solver_text(precision, precond, ctol):
    stop = "Newton stops at max|R| < 1e-8 and max|dx| <= %g max|x| (FP64 norms)" % ctol
    if precision == "f64":
        return "%s; batched CG/%s in FP64 to %g relative residual per Newton step" % (stop, precond, KRYLOV_RTOL)
    return ("%s; FP32 element kernels for J, one batched FP32 CG/%s solve to %g per Newton step (FP32 "
            "V-cycle), FP64 residual: the Newton loop is the FP64 refinement" % (stop, precond, KRYLOV_RTOL))


def _farm_worker(job):
    """One oracle solve (CPU baseline farm): the native build of the oracle's element code."""
    os.environ["ORACLE_BUILD"] = "native"
    from oracle import newton as on
    from tests.oracle_bridge import oracle_problem
    N, sp, lam = job
    op = oracle_problem("bratu", N=N, space=sp, lam=lam)
    on.newton_solve(op, np.zeros(op.ndof), tol=ORACLE_TOL, max_iter=30)
    return 1


def cpu_farm_baseline(specs, order, args):
    """The reference's CPU path timed as its parameter-scan farm (pyoomph/utils/
    paramscan.py:249-277): one single-threaded process per core, each running whole
    Newton solves (oracle element code built -O3 -march=native on this host, vectors_of_pairs
    CSR, SuperLU), over a stratified slice of the sweep, for about --cpu-seconds."""
    import multiprocessing as mp
    import subprocess
    subprocess.check_call(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "native"])
    cores = max(1, min(16, os.cpu_count() or 1))
    k = min(args.cpu_sample, len(order))
    strat = [order[int((j + 0.5) * len(order) / k)] for j in range(k)]
    # a seeded shuffle of the slice, repeated: whatever the time bound cuts off, the solves
    # done are a mix of sizes, not the small end of the size order
    rng = np.random.default_rng(1)
    jobs = [specs[strat[j]] for j in np.concatenate([rng.permutation(len(strat)) for _ in range(4)])]
    env_threads = {v: os.environ.get(v) for v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    for v in env_threads:
        os.environ[v] = "1"
    ctx = mp.get_context("spawn")
    done = 0
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        it = pool.imap_unordered(_farm_worker, jobs)
        for _ in it:
            done += 1
            if time.perf_counter() - t0 > args.cpu_seconds and done >= cores:
                break
        elapsed = time.perf_counter() - t0
        pool.terminate()
    for v, val in env_threads.items():
        if val is None:
            os.environ.pop(v, None)
        else:
            os.environ[v] = val
    return {"value": done / elapsed, "unit": "solves/s", "cores": cores, "kind": "port",
            "host_cores": os.cpu_count(),
            "sample": "%d solves in %.1f s from a stratified slice of the sweep (%d problems, every %d-th by "
                      "unknowns, in seeded random order), %d single-threaded processes (the parameter-scan farm); oracle "
                      "C element code built -O3 -march=native + vectors_of_pairs CSR + scipy SuperLU, "
                      "Newton to max|R| < %g" % (done, elapsed, k, len(order) // k, cores, ORACLE_TOL)}


if __name__ == "__main__":
    main()
