"""Pins the oracle Newton / SuperLU path: manufactured-solution convergence order,
quadratic Newton convergence, and oomph's error paths."""
import numpy as np
import pytest

from oracle import assembly as oa
from oracle import meshes as om
from oracle import newton as on
from tests.oracle_bridge import oracle_problem


def rd_manufactured(N):
    """-lap u + u^3 = f with u = sin(pi x) sin(pi y) sin(pi z), f interpolated at nodes."""
    m = om.cuboid_brick_mesh(N)
    x, y, z = m.coords.T
    u = np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)
    f = 3 * np.pi ** 2 * u + u ** 3
    pinned = m.on_boundary()[:, None]
    op = oa.OracleProblem(m, oa.FORM_RD, 1, pinned, np.zeros((m.nnode, 1)), [], aux=f)
    xs = on.newton_solve(op, np.zeros(op.ndof))
    U = op.nodal_values(xs)[:, 0]
    return np.max(np.abs(U - u))


def test_q2_hex_convergence_order():
    e1, e2 = rd_manufactured(2), rd_manufactured(4)
    assert e2 < e1 / 6.0          # O(h^3) nodal convergence of the Q2 solution


def bratu_centre(N, space):
    op = oracle_problem("bratu", N=N, space=space, lam=1.0)
    xs = on.newton_solve(op, np.zeros(op.ndof))
    U = op.nodal_values(xs)[:, 0]
    c = np.argmin(np.sum((op.mesh.coords - 0.5) ** 2, axis=1))
    return U[c]


def test_bratu_p2_and_p1_agree_on_the_continuum_value():
    a = bratu_centre(16, "C2")
    b = bratu_centre(32, "C1")
    # lambda = 1 on the unit square: u(1/2, 1/2) = 0.0781 (lower branch)
    assert abs(a - 0.0781) < 5e-4 and abs(b - a) < 2e-4


def test_newton_converges_quadratically():
    op = oracle_problem("bratu", N=8, space="C2", lam=2.5)
    hist = []
    on.newton_solve(op, np.zeros(op.ndof), tol=1e-13, history=hist)
    assert len(hist) >= 3
    assert hist[2] < hist[1] ** 1.5 * 10


def test_always_one_step_on_converged_start():
    op = oracle_problem("poisson", N=4)
    x = on.newton_solve(op, np.zeros(op.ndof))
    hist = []
    on.newton_solve(op, x, history=hist)
    assert len(hist) == 2          # initial residual + one step taken anyway


def test_max_iterations_raises():
    op = oracle_problem("bratu", N=6, space="C2", lam=3.0)
    with pytest.raises(on.NewtonSolverError):
        on.newton_solve(op, np.zeros(op.ndof), max_iter=1, tol=1e-14)


def test_nan_is_never_converged():
    assert on.max_abs(np.array([0.0, np.nan])) == float("inf")
    op = oracle_problem("bratu", N=4, space="C1", lam=1.0)
    x0 = np.zeros(op.ndof)
    x0[0] = np.nan
    # NaN contributions are dropped from the CSR (fabs(NaN) > 0 is false), so SuperLU
    # reports a singular factor -- a RuntimeError, as in pyoomph/solvers/scipy.py:68-89
    with pytest.raises(RuntimeError):
        on.newton_solve(op, x0)


def test_superlu_transpose_convention():
    rng = np.random.default_rng(0)
    A = rng.normal(size=(5, 5)) + 5 * np.eye(5)
    A[A < -1.0] = 0.0
    import scipy.sparse as sp
    S = sp.csr_matrix(A)
    b = rng.normal(size=5)
    x = on.superlu_solve(5, S.indptr, S.indices, S.data, b)
    assert np.allclose(A @ x, b)


def test_superlu_solve_leaves_inputs_untouched():
    op = oracle_problem("poisson", N=3)
    r, rs, col, val = op.get_jacobian(np.zeros(op.ndof))
    c0, v0 = col.copy(), val.copy()
    on.superlu_solve(op.ndof, rs, col, val, r)
    assert np.array_equal(col, c0) and np.array_equal(val, v0)


def test_converged_solve_reaches_roundoff_where_the_absolute_test_stops_early():
    """A small source (residual scale 1e-6) on the Q2 cube: the absolute test max|R| < 1e-12
    accepts the first Newton iterate of the cubic problem, whose error is still ~1e-6 of the
    solution; one more step (converged_solve) lands on the fixed point: a further Newton
    step moves it by roundoff only."""
    m = om.cuboid_brick_mesh(4)
    rng = np.random.default_rng(7)
    f = 40.0 * rng.uniform(-1.0, 1.0, m.nnode)
    op = oa.OracleProblem(m, oa.FORM_RD, 1, m.on_boundary()[:, None], np.zeros((m.nnode, 1)), [], aux=f)
    hist = []
    x1 = on.newton_solve(op, np.zeros(op.ndof), tol=1e-6, history=hist)
    xc = on.converged_solve(op, np.zeros(op.ndof), tol=1e-6)
    x_next = xc - on.superlu_solve(op.ndof, *op.get_jacobian(xc)[1:], op.get_residuals(xc))
    scale = np.max(np.abs(xc))
    assert np.max(np.abs(x_next - xc)) <= 1e-14 * scale
    assert np.max(np.abs(x1 - xc)) > 1e3 * np.max(np.abs(x_next - xc))
