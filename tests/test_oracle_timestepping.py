"""Pins the oracle's time derivative, mass matrix and BDF time stepping (no GPU).

* BDF weights: BDF1 differentiates linear functions exactly, BDF2 quadratics, for
  unequal steps (a property of the formulas, not a retyping of them);
* element mass matrices (flag 2) against closed forms: P1 (area/12 [2 1 1; 1 2 1; 1 1 2]),
  Q2 quads and Q2 bricks (1D Lagrange mass tensor products), the P2 mass matrix via
  u^T M u = int u^2 for linear u;
* the time term of the residual equals M (du/dt at the nodes), and the transient
  Jacobian matches finite differences of the residual with the history held (scalar
  forms and Taylor-Hood, whose mass matrix is rho psi_j psi_i on the velocity blocks);
* time convergence of the whole stepping loop on the heat equation against the exact
  semi-discrete solution expm(-t M^-1 K) u0: first order for BDF1, second order for BDF2
  with its BDF1 first step.
"""
import numpy as np
import pytest
import scipy.linalg
import scipy.sparse

from oracle import assembly as oa
from oracle import meshes as om
from oracle import timestepping as ots
from tests.oracle_bridge import oracle_problem
from tests.test_oracle_elements import _fd_jacobian, _lag_1d, free_problem


def element_mass(op, e):
    x = np.zeros(op.ndof)
    op.set_time(ots.bdf1_weights(1.0), x, x)
    try:
        _, M = op.element_outputs(x, 2)
    finally:
        op.set_time(None, None, None)
    nv = op.nvar[e]
    return M[e].reshape(-1)[:nv * nv].reshape(nv, nv)


def dense(rs, col, val, n):
    return scipy.sparse.csr_matrix((val, col, rs), shape=(n, n)).toarray()


def test_bdf_weights_differentiate_polynomials_exactly():
    dt, dtp, t = 0.3, 0.7, 2.0
    w1 = ots.bdf1_weights(dt)
    for a, b in ((1.0, 2.0), (-3.0, 0.5)):
        f = lambda s: a + b * s
        assert abs(w1[0] * f(t) + w1[1] * f(t - dt) - b) < 1e-12
    w2 = ots.bdf2_weights(dt, dtp)
    for c in ((1.0, 2.0, 3.0), (0.0, -1.0, 4.0)):
        f = lambda s: c[0] + c[1] * s + c[2] * s * s
        d = c[1] + 2 * c[2] * t
        assert abs(w2[0] * f(t) + w2[1] * f(t - dt) + w2[2] * f(t - dt - dtp) - d) < 1e-11
    assert abs(w2.sum()) < 1e-12


def test_p1_mass_closed_form():
    m = om.rectangular_quad_mesh(1, "left", "C1")
    op = free_problem(m, oa.FORM_POISSON, [])
    for e in range(m.nelement):
        X = m.coords[m.elements[e]]
        area = abs(np.linalg.det(np.array([X[0] - X[2], X[1] - X[2]]))) / 2
        M = area / 12 * np.array([[2, 1, 1], [1, 2, 1], [1, 1, 2]])
        # the TGauss<2,2> knots and weights are decimal literals (3e-13 relative)
        assert np.allclose(element_mass(op, e), M, rtol=1e-12, atol=0)


def test_p2_mass_integrates_squares_of_linear_functions():
    m = om.rectangular_quad_mesh(2, "left", "C2")
    m.coords[:, 0] *= 1.3
    op = free_problem(m, oa.FORM_POISSON, [])
    for e in range(m.nelement):
        X = m.coords[m.elements[e]]
        M = element_mass(op, e)
        v = X[:3]
        area = abs(np.linalg.det(np.array([v[0] - v[2], v[1] - v[2]]))) / 2
        for a, b, c in ((1.0, 0.0, 0.0), (0.3, 1.0, -2.0)):
            u = a + b * X[:, 0] + c * X[:, 1]
            # int of a quadratic over a triangle: the 3 edge-midpoint rule is exact
            mids = [(v[0] + v[1]) / 2, (v[1] + v[2]) / 2, (v[2] + v[0]) / 2]
            exact = area / 3 * sum((a + b * q[0] + c * q[1]) ** 2 for q in mids)
            assert abs(u @ M @ u - exact) < 1e-12 * max(1.0, exact)


def test_q2_and_brick_mass_tensor_products():
    m = om.rectangular_quad_mesh(1, False, "C2")
    m.coords[:, 1] *= 0.5
    op = free_problem(m, oa.FORM_POISSON, [])
    _, Mx = _lag_1d(1.0)
    _, My = _lag_1d(0.5)
    # Gauss<2,3>'s mistyped knot makes the rule inexact by ~1e-8 relative
    assert np.allclose(element_mass(op, 0), np.kron(My, Mx), rtol=0, atol=5e-9)
    b = om.cuboid_brick_mesh(1)
    opb = free_problem(b, oa.FORM_RD, [], aux=np.zeros(b.nnode))
    _, M1 = _lag_1d(1.0)
    assert np.allclose(element_mass(opb, 0), np.kron(np.kron(M1, M1), M1), rtol=0, atol=1e-14)


@pytest.mark.parametrize("kind,kw", [("bratu", dict(N=3, space="C2", lam=2.0)),
                                     ("poisson", dict(N=3)), ("ns", dict(N=3, Re=50.0))])
def test_time_term_is_mass_times_nodal_rate(kind, kw):
    op = oracle_problem(kind, **kw)
    rng = np.random.default_rng(4)
    x, h1, h2 = (rng.uniform(-0.5, 0.5, op.ndof) for _ in range(3))
    tw = ots.bdf2_weights(0.1, 0.2)
    r_steady = op.get_residuals(x)
    op.set_time(tw, h1, h2)
    r_tr = op.get_residuals(x)
    _, rs, col, val = op.get_mass_matrix(x)
    op.set_time(None, None, None)
    M = dense(rs, col, val, op.ndof)
    rate = tw[0] * x + tw[1] * h1 + tw[2] * h2      # pinned nodes: weights sum to 0
    assert np.allclose(r_tr - r_steady, M @ rate, rtol=0, atol=1e-13)


@pytest.mark.parametrize("kind,kw", [("bratu", dict(N=2, space="C2", lam=1.5)), ("ns", dict(N=2))])
def test_transient_jacobian_matches_finite_differences(kind, kw):
    op = oracle_problem(kind, **kw)
    rng = np.random.default_rng(6)
    x, h1, h2 = (rng.uniform(-0.5, 0.5, op.ndof) for _ in range(3))
    op.set_time(ots.bdf2_weights(0.05, 0.05), h1, h2)
    _, rs, col, val = op.get_jacobian(x)
    J = dense(rs, col, val, op.ndof)
    Jfd = _fd_jacobian(op, x)
    op.set_time(None, None, None)
    assert np.max(np.abs(J - Jfd)) < 1e-7 * max(1.0, np.max(np.abs(J)))


def _heat_problem(N):
    op = oracle_problem("poisson", N=N)
    op.params = np.zeros(0)        # no source: du/dt = lap u
    return op


def _heat_errors(bdf2, steps_list, T=0.05):
    op = _heat_problem(6)
    xy = op.mesh.coords
    free = op.eqn[:, 0] >= 0
    u0 = np.zeros(op.ndof)
    u0[op.eqn[free, 0]] = (np.sin(np.pi * xy[free, 0]) * np.sin(np.pi * xy[free, 1]))
    _, rs, col, val = op.get_jacobian(np.zeros(op.ndof))
    K = dense(rs, col, val, op.ndof)
    op.set_time(ots.bdf1_weights(1.0), u0, u0)
    _, rs, col, val = op.get_mass_matrix(u0)
    op.set_time(None, None, None)
    M = dense(rs, col, val, op.ndof)
    exact = scipy.linalg.expm(-T * np.linalg.solve(M, K)) @ u0
    errs = []
    for n in steps_list:
        dt = T / n
        st = ots.TimeState(u0)
        x = u0.copy()
        for _ in range(n):
            if not bdf2:
                st.steps_done = 0   # BDF1 throughout
            x = ots.unsteady_newton_step(op, x, st, dt, tol=1e-13)
        errs.append(np.max(np.abs(x - exact)))
    return errs


def test_bdf1_converges_first_order():
    e = _heat_errors(False, [8, 16, 32])
    assert 1.8 < e[0] / e[1] < 2.2 and 1.8 < e[1] / e[2] < 2.2


def test_bdf2_converges_second_order():
    e = _heat_errors(True, [8, 16, 32])
    assert 3.4 < e[0] / e[1] < 4.6 and 3.4 < e[1] / e[2] < 4.6
