"""The benchmark configurations on the GPU, each against the oracle or a known answer.

* Batched Bratu sweep: a stratified slice of the 512-problem sweep (N ~ U{16..128}, P1 or
  P2, lambda ~ U(0.5, 3)) -- the largest P1 and P2 problems, the smallest and largest
  lambda, and the rest spread over the sizes -- solved as ONE batch with the benchmark's
  solver (CG + smoothed-aggregation multigrid, FP64), each sub-problem within 1e-8 of the
  oracle's SuperLU Newton.
* 3D reaction-diffusion with its configured mixed path: FP32 element kernels for the
  Jacobian and FP32 ILU(0)-GMRES inside FP64 iterative refinement, FP64 residual; Newton to
  max|R| < 1e-13, within 1e-8 of the oracle.
* Lid-driven cavity: 64 x 64 against the oracle (1e-8) and 128 x 128 against the
  centreline velocities of Ghia, Ghia & Shin (1982) and the converged extrema.
* Solver regression cases: the GMRES restart of a solve does not depend on the restart of
  an earlier solve on the same problem (bitwise); GMRES with the multigrid preconditioner in
  a batch whose sub-problems stop at different iterations equals each sub-problem alone
  (bitwise) and the oracle; continuation with GMRES + multigrid on the lower branch equals
  the oracle and the batch equals alone; a multigrid set-up on an indefinite Jacobian
  reports the failure instead of iterating with an invalid V-cycle.
"""
import numpy as np
import pytest

from oracle import newton as on
from pyoomph_amd import workloads as wl
from tests.oracle_bridge import oracle_problem, product_problem

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def sweep_slice(n=16):
    """Indices into the seed-0 sweep: largest P2, largest P1, smallest and largest lambda,
    then every k-th problem by unknowns."""
    specs = wl.bratu_sweep_specs(512, 0)
    dofs = [(2 * N - 1) ** 2 if sp == "C2" else (N - 1) ** 2 for N, sp, _ in specs]
    pick = []
    for sp in ("C2", "C1"):
        cand = [i for i in range(512) if specs[i][1] == sp]
        pick.append(max(cand, key=lambda i: (specs[i][0], i)))
    pick.append(min(range(512), key=lambda i: specs[i][2]))
    pick.append(max(range(512), key=lambda i: specs[i][2]))
    order = sorted(range(512), key=lambda i: dofs[i])
    k = 0
    while len(set(pick)) < n:
        pick.append(order[(k * 37) % 512])
        k += 1
    return specs, list(dict.fromkeys(pick))


def test_bratu_sweep_slice_batch_matches_oracle(gpu):
    from pyoomph_amd.problem import Problem
    specs, idx = sweep_slice(16)
    sl = [specs[i] for i in idx]
    assert any(s[0] >= 120 and s[1] == "C2" for s in sl) and any(s[0] >= 120 and s[1] == "C1" for s in sl)
    assert min(s[2] for s in sl) < 0.52 and max(s[2] for s in sl) > 2.98
    P = Problem([wl.bratu(*s) for s in sl])
    P.krylov, P.preconditioner = "cg", "amg"
    st = P.solve()
    x = P.get_current_dofs()
    for i, (N, sp, lam) in enumerate(sl):
        print("sweep problem", idx[i], N, sp, "%.4f" % lam, flush=True)
        assert st[i]["converged"] and st[i]["linear_failures"] == 0
        op = oracle_problem("bratu", N=N, space=sp, lam=lam)
        xo = on.newton_solve(op, np.zeros(op.ndof))
        assert rel(P.sub_dofs(i, x), xo) <= 1e-8, (idx[i], N, sp, lam)


def test_multigrid_batch_with_empty_prolongation_blocks(gpu):
    """Forty tiny problems between two larger ones: their hierarchies end on the finest
    level, so the finest prolongation has 256-row blocks without entries, starting at an
    odd entry (checked on the host plan in test_amg_plan). Every spmv variant runs on that
    matrix when the level's launches are tuned (a variant that read entries of such a block
    would fault); the batch still matches the oracle."""
    from pyoomph_amd.problem import Problem
    specs = [(40, "C1", 1.0)] + [(7, "C1", 1.0 + 0.01 * i) for i in range(40)] + [(40, "C2", 1.2)]
    P = Problem([wl.bratu(*s) for s in specs])
    P.krylov, P.preconditioner = "cg", "amg"
    st = P.solve()
    x = P.get_current_dofs()
    for i, (N, sp, lam) in enumerate(specs):
        assert st[i]["converged"] and st[i]["linear_failures"] == 0
        if i in (0, 1, 20, 40, 41):
            op = oracle_problem("bratu", N=N, space=sp, lam=lam)
            assert rel(P.sub_dofs(i, x), on.newton_solve(op, np.zeros(op.ndof))) <= 1e-8, (i, N, sp)


def test_reaction_diffusion_mixed_path_matches_oracle(gpu):
    from pyoomph_amd.problem import Problem
    N = 10
    mp = product_problem("rd", N=N, seed=0)
    op = oracle_problem("rd", N=N, f=mp.equations.source_field)
    P = Problem(mp)
    P.krylov, P.preconditioner = "gmres", "ilu0"
    P.jacobian_precision, P.linear_precision = "f32", "f32"
    # an FP32 Jacobian converges linearly: iterate the FP64 residual well below 1e-8 so
    # that the solution, not only the residual, is within 1e-8
    P.newton_solver_tolerance = 1e-13
    P.max_newton_iterations = 20
    st = P.solve()
    assert st[0]["converged"]
    xo = on.newton_solve(op, np.zeros(op.ndof), tol=1e-13, max_iter=20)
    assert rel(P.get_current_dofs(), xo) <= 1e-8


def test_cavity_64_matches_oracle(gpu):
    from pyoomph_amd.problem import Problem
    P = Problem(product_problem("ns", N=64))
    P.krylov, P.preconditioner, P.gmres_restart = "gmres", "ilu0", 300
    # the Krylov solves stop at 1e-10 relative (inexact Newton): iterate the residual to
    # 1e-12 so that the solution, not only the residual, is within 1e-8 of the oracle's
    P.gmres_max_iters, P.newton_solver_tolerance = 8000, 1e-12
    P.max_newton_iterations = 15
    st = P.solve()
    assert st[0]["converged"], st
    op = oracle_problem("ns", N=64)
    xo = on.newton_solve(op, np.zeros(op.ndof), tol=1e-12, max_iter=15)
    assert rel(P.get_current_dofs(), xo) <= 1e-8


def test_cavity_128_matches_ghia(gpu):
    from pyoomph_amd.problem import Problem
    from tests.test_oracle_ns_known_answers import GHIA_U, GHIA_V
    mp = product_problem("ns", N=128)
    P = Problem(mp)
    P.krylov, P.preconditioner, P.gmres_restart = "gmres", "ilu0", 150
    P.gmres_max_iters, P.newton_solver_tolerance = 8000, 1e-10
    st = P.solve()
    assert st[0]["converged"], st
    U = mp.nodal_values(P.get_current_dofs())
    c = mp.mesh.coords

    def at(px, py):
        d = np.abs(c[:, 0] - px) + np.abs(c[:, 1] - py)
        i = int(np.argmin(d))
        assert d[i] < 1e-12
        return i

    # nodes 1/256 apart: Ghia's points k/128 are nodes
    uc = {k: U[at(0.5, k / 128), 0] for k in range(129)}
    vc = {k: U[at(k / 128, 0.5), 1] for k in range(129)}
    assert max(abs(uc[k] - u) for k, u in GHIA_U) < 1e-2
    assert max(abs(vc[k] - v) for k, v in GHIA_V) < 1e-2
    assert abs(min(uc.values()) - (-0.2140)) < 1.5e-3
    assert abs(max(vc.values()) - 0.1796) < 1.5e-3
    assert abs(min(vc.values()) - (-0.2538)) < 1.5e-3


def _gmres_problem(restart):
    from pyoomph_amd.problem import Problem
    P = Problem(product_problem("ns", N=8))
    P.krylov, P.preconditioner, P.gmres_restart = "gmres", "ilu0", restart
    P.newton_solver_tolerance = 1e-10
    return P


def test_gmres_restart_is_the_requested_one(gpu):
    """A solve with restart 50 after one with restart 150 on the same problem gives the
    bits of a fresh restart-50 problem (the workspace stays sized for 150)."""
    P = _gmres_problem(150)
    P.solve()
    P.dofs.zero_()
    P.gmres_restart = 50
    st = P.solve()
    fresh = _gmres_problem(50)
    st_f = fresh.solve()
    assert st[0]["linear_iters"] == st_f[0]["linear_iters"]
    assert np.array_equal(P.get_current_dofs(), fresh.get_current_dofs())


GMRES_AMG_SPECS = [(12, "C2", 1.5), (20, "C1", 2.8), (9, "C2", 0.6), (15, "C1", 0.9)]


def test_gmres_multigrid_batch_equals_alone_and_oracle(gpu):
    """GMRES + multigrid filters the V-cycle and the spmv by the sub-problems still
    iterating: in a batch whose sub-problems stop at different iterations each one gets the
    bits it gets alone, and the oracle's solution."""
    from pyoomph_amd.problem import Problem

    def run(specs):
        P = Problem([wl.bratu(*s) for s in specs])
        P.krylov, P.preconditioner, P.gmres_restart = "gmres", "amg", 60
        st = P.solve()
        return P, st

    B, stb = run(GMRES_AMG_SPECS)
    xb = B.get_current_dofs()
    iters = set()
    for i, s in enumerate(GMRES_AMG_SPECS):
        A, sta = run([s])
        iters.add(sta[0]["linear_iters"])
        assert np.array_equal(A.get_current_dofs(), B.sub_dofs(i, xb)), s
        op = oracle_problem("bratu", N=s[0], space=s[1], lam=s[2])
        xo = on.newton_solve(op, np.zeros(op.ndof))
        assert rel(B.sub_dofs(i, xb), xo) <= 1e-8, s
    assert len(iters) > 1   # the sub-problems did stop at different iterations


def test_continuation_multigrid_lower_branch(gpu):
    """Arclength steps with GMRES + multigrid below the fold: batch equals alone bitwise and
    every point solves the oracle's equations."""
    from oracle.newton import max_abs
    from tests.test_gpu_continuation import SPECS, _gpu_run
    both = _gpu_run(SPECS, 4, precond="amg")
    for i, sp in enumerate(SPECS):
        alone = _gpu_run([sp], 4, precond="amg")
        for k in range(4):
            assert both[k][0][i] == alone[k][0][0]
            assert np.array_equal(both[k][1][i], alone[k][1][0])
            lam = both[k][0][i]
            assert lam < 6.0            # still on the lower branch
            chk = oracle_problem("bratu", lam=lam, **sp)
            assert max_abs(chk.get_residuals(both[k][1][i])) < 1e-8


def test_multigrid_on_indefinite_jacobian_reports_failure(gpu):
    """Far up the Bratu branch (lambda e^u dominates the Laplacian) the Jacobian is
    indefinite; the multigrid set-up must say so rather than precondition with an
    indefinite coarsest factor."""
    from pyoomph_amd.problem import Problem
    mp = wl.bratu(12, "C2", 3.0)
    P = Problem(mp)
    P.krylov, P.preconditioner = "gmres", "amg"
    xy = mp.mesh.coords
    u = 6.0 * np.sin(np.pi * xy[:, 0]) * np.sin(np.pi * xy[:, 1])
    x0 = np.zeros(P.ndof)
    m = mp.eqn[:, 0] >= 0
    x0[mp.eqn[m, 0]] = u[m]
    P.set_current_dofs(x0)
    with pytest.raises(RuntimeError, match="not positive definite"):
        P.solve()


@pytest.fixture(scope="module")
def converged_slice():
    """The slice and the oracle's converged solutions (SuperLU Newton to max|R| < 1e-13)."""
    specs, idx = sweep_slice(16)
    sl = [specs[i] for i in idx]
    ref = []
    for N, sp, lam in sl:
        op = oracle_problem("bratu", N=N, space=sp, lam=lam)
        ref.append(on.converged_solve(op, np.zeros(op.ndof), tol=1e-13))
    return idx, sl, ref


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_benchmark_settings_reach_the_converged_solution(gpu, converged_slice, precision):
    """The benchmark's solver settings on a slice of the sweep: inexact Newton (CG to 5e-3
    relative residual per step; in mixed precision FP32 element kernels for J and one FP32
    CG solve with the FP32 V-cycle per step, FP64 residual) stopping at max|R| < 1e-8 and
    an FP64 update below 1e-6 of the solution. Every sub-problem is within 1e-8 of the
    converged FP64 oracle solution (pyoomph's residual test alone is not: on 128 x 128 P2
    meshes it stops ~1e-6 away)."""
    from pyoomph_amd.problem import Problem
    idx, sl, ref = converged_slice
    P = Problem([wl.bratu(*s) for s in sl])
    P.krylov, P.preconditioner = "cg", "amg"
    P.jacobian_precision = P.linear_precision = "f32" if precision == "mixed" else "f64"
    P.gmres_rtol = P.rtol_f32 = 5e-3
    P.max_refinements = 1
    P.newton_correction_tolerance = 1e-6
    P.max_newton_iterations = 20
    st = P.solve()
    x = P.get_current_dofs()
    for i, (N, sp, lam) in enumerate(sl):
        print(precision, "sweep problem", idx[i], N, sp, "%.4f" % lam, st[i]["newton_iters"], flush=True)
        assert st[i]["converged"]
        assert rel(P.sub_dofs(i, x), ref[i]) <= 1e-8, (idx[i], N, sp, lam)


@pytest.mark.parametrize("kind,kw", [("bratu", dict(N=64, space="C2", lam=2.5)), ("rd", dict(N=6, seed=1))])
def test_multigrid_element_level(gpu, kind, kw):
    """The multigrid's element level (P2 triangles, Q2 bricks coarsen first to the linear
    functions on their vertices): the solve reaches the oracle's solution with and without
    it, and with it the conjugate-gradient solves take fewer iterations: at most 70 % on P2
    triangles, whose stencil aggregates poorly (bricks: 10 against 13 at N = 6)."""
    from pyoomph_amd.problem import Problem
    mp = product_problem(kind, **kw)
    if kind == "rd":
        op = oracle_problem("rd", N=kw["N"], f=mp.equations.source_field)
    else:
        op = oracle_problem(kind, **kw)
    xo = on.newton_solve(op, np.zeros(op.ndof))
    iters = {}
    for level in (True, False):
        P = Problem(mp)
        P.krylov, P.preconditioner = "cg", "amg"
        P.amg_element_level = level
        st = P.solve()
        print(kind, "element level", level, "linear iterations", st[0]["linear_iters"], flush=True)
        assert st[0]["converged"] and st[0]["linear_failures"] == 0
        assert rel(P.get_current_dofs(), xo) <= 1e-8
        iters[level] = st[0]["linear_iters"]
    assert iters[True] <= (0.7 * iters[False] if kind == "bratu" else iters[False] - 1), iters


def test_multigrid_element_level_switch_rebuilds_the_hierarchy(gpu):
    """Turning the element level off and on again on one problem gives the bits of fresh
    problems with each setting."""
    from pyoomph_amd.problem import Problem
    specs = [(24, "C2", 1.5), (20, "C1", 2.0)]

    def solve(P, level):
        P.krylov, P.preconditioner = "cg", "amg"
        P.amg_element_level = level
        P.dofs.zero_()
        P.solve()
        return P.get_current_dofs()

    P = Problem([wl.bratu(*s) for s in specs])
    a_on, a_off, a_on2 = solve(P, True), solve(P, False), solve(P, True)
    f_on = solve(Problem([wl.bratu(*s) for s in specs]), True)
    f_off = solve(Problem([wl.bratu(*s) for s in specs]), False)
    assert np.array_equal(a_on, f_on) and np.array_equal(a_on2, f_on) and np.array_equal(a_off, f_off)
    assert not np.array_equal(a_on, a_off)
