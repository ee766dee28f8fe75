"""Pins the oracle's Taylor-Hood Navier-Stokes residual with answers that do not come from
the oracle itself.

The reference's weak form (pyoomph/equations/navier_stokes.py:303-347, 481-503):
  momentum, test v:   weak(sigma, grad v) + weak(rho * material_derivative(u, u), v)
                      sigma = 2 mu sym(grad u) - p I,  (u . grad) u in component a is
                      sum_b u_b d_b u_a
  continuity, test q: weak(div u, q)
with Taylor-Hood Q2 velocity / Q1 pressure (navier_stokes.py:255-264).

* Element known answer: on one affinely distorted Q2/Q1 element, for a quadratic velocity
  with a non-symmetric gradient and a linear pressure (both represented exactly by the
  element), the oracle's element residual equals the integrals above computed here with
  an 8x8 Gauss-Legendre rule and Lagrange bases written out in this file. The test also
  shows that a transposed convective index, a flipped continuity sign and a dropped rho
  each move the answer far outside the tolerance.
* Poiseuille: u = (y (1 - y), 0), p = -2 mu x + c solves the steady equations exactly and
  lies in the discrete space, so every free row of the oracle's residual vanishes.
* Lid-driven cavity at Re = 100 against the centreline velocities of Ghia, Ghia & Shin
  (J. Comput. Phys. 48 (1982) 387-411, Tables I and II) to 1e-2, and the extrema of the
  converged solution to 1.5e-3. On 64 x 64 Q2 elements the tabulated points (multiples
  of 1/128) are nodes, so no interpolation enters.

Tolerances: the oracle integrates with the reference's Gauss<2,3>, whose knot
0.774596662941483 is mistyped in oomph (integral.cc), so exact integrals agree to about
1e-8 relative, not to rounding; the tests use 1e-7.
"""
import numpy as np
import pytest

from oracle import assembly as oa
from oracle import meshes as om
from oracle import newton as on
from tests.oracle_bridge import oracle_problem

# Q2 local node l = 3*iy + ix at reference position (s0, s1) = (ix - 1, iy - 1);
# pressure (Q1) on the corner nodes 0, 2, 6, 8 in that order.
CORNERS = [0, 2, 6, 8]
# oomph's local dof order of the Taylor-Hood element: per node ux, uy (and p on corners)
SLOTS = []
for _l in range(9):
    SLOTS += [(_l, 0), (_l, 1)] + ([(_l, 2)] if _l in CORNERS else [])


def lag1(s):
    """1D quadratic Lagrange basis at -1, 0, 1 and its derivative."""
    return (np.array([0.5 * s * (s - 1), 1 - s * s, 0.5 * s * (s + 1)]),
            np.array([s - 0.5, -2 * s, s + 0.5]))


def q2_basis(s0, s1):
    L0, D0 = lag1(s0)
    L1, D1 = lag1(s1)
    psi = np.array([L0[l % 3] * L1[l // 3] for l in range(9)])
    dpsi = np.array([[D0[l % 3] * L1[l // 3], L0[l % 3] * D1[l // 3]] for l in range(9)])
    return psi, dpsi


def q1_basis(s0, s1):
    n0 = np.array([(1 - s0) / 2, (1 + s0) / 2])
    n1 = np.array([(1 - s1) / 2, (1 + s1) / 2])
    return np.array([n0[0] * n1[0], n0[1] * n1[0], n0[0] * n1[1], n0[1] * n1[1]])


def element_residual_exact(X, ufun, gradu, pfun, mu, rho, mistake=None):
    """Residual of one affine Q2/Q1 element in oomph's slot order, by 8x8 Gauss-Legendre.
    X: [9, 2] node positions (affine image of the reference square)."""
    c = X[4]
    M = np.column_stack([(X[2] - X[0]) / 2, (X[6] - X[0]) / 2])   # x = c + M s
    det = abs(np.linalg.det(M))
    Minv_T = np.linalg.inv(M).T
    g, wg = np.polynomial.legendre.leggauss(8)
    R = np.zeros(len(SLOTS))
    for a, wa in zip(g, wg):
        for b, wb in zip(g, wg):
            x = c + M @ np.array([a, b])
            psi, dpsi = q2_basis(a, b)
            dx = dpsi @ Minv_T.T          # [9, 2] physical gradients
            C = q1_basis(a, b)
            u = ufun(x)
            G = gradu(x)                   # G[a][b] = d u_a / d x_b
            p = pfun(x)
            sig = mu * (G + G.T) - p * np.eye(2)
            conv = G @ u                   # sum_b u_b d_b u_a
            if mistake == "transposed_convection":
                conv = G.T @ u
            r = rho if mistake != "no_rho" else 1.0
            div = G[0, 0] + G[1, 1]
            if mistake == "continuity_sign":
                div = -div
            W = wa * wb * det
            for k, (l, v) in enumerate(SLOTS):
                if v < 2:
                    R[k] += W * (sig[v] @ dx[l] + r * conv[v] * psi[l])
                else:
                    R[k] += W * div * C[CORNERS.index(l)]
    return R


def _single_element(X):
    m = om.rectangular_quad_mesh(1, False, "C2")
    m.coords = np.asarray(X, dtype=np.float64).copy()
    return m


def test_single_element_residual_known_answer():
    # unit square mapped by a non-symmetric affine map (no axis alignment, no symmetry)
    m0 = om.rectangular_quad_mesh(1, False, "C2")
    A = np.array([[1.3, 0.4], [-0.2, 0.9]])
    X = m0.coords @ A.T + np.array([0.25, -0.6])
    m = _single_element(X)

    def ufun(x):
        X_, Y_ = x
        return np.array([0.3 + 0.5 * X_ - 0.7 * Y_ + 0.2 * X_ * X_ - 0.4 * X_ * Y_ + 0.6 * Y_ * Y_,
                         -0.1 + 0.8 * X_ + 0.3 * Y_ - 0.5 * X_ * X_ + 0.25 * X_ * Y_ - 0.35 * Y_ * Y_])

    def gradu(x):
        X_, Y_ = x
        return np.array([[0.5 + 0.4 * X_ - 0.4 * Y_, -0.7 - 0.4 * X_ + 1.2 * Y_],
                         [0.8 - 1.0 * X_ + 0.25 * Y_, 0.3 + 0.25 * X_ - 0.7 * Y_]])

    def pfun(x):
        return 0.4 - 0.9 * x[0] + 1.1 * x[1]

    mu, rho = 0.3, 1.7
    nn = m.nnode
    pinned = np.zeros((nn, 3), bool)
    corner_nodes = set(m.elements[0, CORNERS].tolist())
    for n in range(nn):
        if n not in corner_nodes:
            pinned[n, 2] = True          # pressure dummies on non-corner nodes
    bc = np.zeros((nn, 3))
    op = oa.OracleProblem(m, oa.FORM_NS, 3, pinned, bc, [mu, rho])
    vals = np.zeros((nn, 3))
    for n in range(nn):
        vals[n, :2] = ufun(m.coords[n])
        vals[n, 2] = pfun(m.coords[n])
    x = np.zeros(op.ndof)
    free = op.eqn >= 0
    x[op.eqn[free]] = vals[free]
    R, _ = op.element_outputs(x, 0)
    assert op.nvar[0] == 22          # every slot free: local order = slot order
    r_oracle = R[0, :22]
    Xe = m.coords[m.elements[0]]
    r_exact = element_residual_exact(Xe, ufun, gradu, pfun, mu, rho)
    scale = np.max(np.abs(r_exact))
    assert np.max(np.abs(r_oracle - r_exact)) <= 1e-7 * scale
    # the check discriminates: each plausible mistake is far outside the tolerance
    for mistake in ("transposed_convection", "continuity_sign", "no_rho"):
        wrong = element_residual_exact(Xe, ufun, gradu, pfun, mu, rho, mistake)
        assert np.max(np.abs(r_oracle - wrong)) > 1e-2 * scale, mistake


@pytest.mark.parametrize("N,mu,rho", [(4, 0.05, 1.3), (3, 1.0, 0.4)])
def test_poiseuille_is_an_exact_discrete_solution(N, mu, rho):
    m = om.rectangular_quad_mesh(N, False, "C2")
    nn = m.nnode
    Xc, Yc = m.coords[:, 0], m.coords[:, 1]
    exact = np.column_stack([Yc * (1 - Yc), np.zeros(nn), -2 * mu * Xc + 0.7])
    vert = np.zeros(nn, bool)
    vert[m.elements[:, CORNERS].ravel()] = True
    bnd = m.on_boundary({"left", "right", "bottom", "top"})
    pinned = np.zeros((nn, 3), bool)
    pinned[bnd, 0] = pinned[bnd, 1] = True
    pinned[~vert, 2] = True
    pinned[0, 2] = True
    op = oa.OracleProblem(m, oa.FORM_NS, 3, pinned, exact * pinned, [mu, rho])
    x = np.zeros(op.ndof)
    free = op.eqn >= 0
    x[op.eqn[free]] = exact[free]
    r = op.get_residuals(x)
    # scale of the individual terms: viscous stress mu * |u'| and the pressure
    assert np.max(np.abs(r)) <= 1e-7 * max(mu, 1.0)
    # a wrong pressure sign is not a solution
    exact_bad = exact.copy()
    exact_bad[:, 2] = -exact_bad[:, 2]
    xb = np.zeros(op.ndof)
    xb[op.eqn[free]] = exact_bad[free]
    assert np.max(np.abs(op.get_residuals(xb))) > 1e-3


# Ghia, Ghia & Shin (1982), Re = 100: u on the vertical centreline x = 0.5 (Table I) and
# v on the horizontal centreline y = 0.5 (Table II); positions as k/128.
GHIA_U = [(128, 1.0), (125, 0.84123), (124, 0.78871), (123, 0.73722), (122, 0.68717),
          (109, 0.23151), (94, 0.00332), (79, -0.13641), (64, -0.20581), (58, -0.21090),
          (36, -0.15662), (22, -0.10150), (13, -0.06434), (9, -0.04775), (8, -0.04192),
          (7, -0.03717), (0, 0.0)]
GHIA_V = [(128, 0.0), (124, -0.05906), (123, -0.07391), (122, -0.08864), (121, -0.10313),
          (116, -0.16914), (110, -0.22445), (103, -0.24533), (64, 0.05454), (30, 0.17527),
          (29, 0.17507), (20, 0.16077), (12, 0.12317), (10, 0.10890), (9, 0.10091),
          (8, 0.09233), (0, 0.0)]


@pytest.fixture(scope="module")
def cavity64():
    op = oracle_problem("ns", N=64)
    x = on.newton_solve(op, np.zeros(op.ndof), tol=1e-10)
    return op, op.nodal_values(x)


def _node_at(op, px, py):
    c = op.mesh.coords
    d = np.abs(c[:, 0] - px) + np.abs(c[:, 1] - py)
    i = int(np.argmin(d))
    assert d[i] < 1e-12
    return i


def test_cavity_re100_matches_ghia_centrelines(cavity64):
    op, U = cavity64
    uc = {k: U[_node_at(op, 0.5, k / 128), 0] for k in range(129)}
    vc = {k: U[_node_at(op, k / 128, 0.5), 1] for k in range(129)}
    du = [abs(uc[k] - u) for k, u in GHIA_U]
    dv = [abs(vc[k] - v) for k, v in GHIA_V]
    # Ghia's table is a second-order finite-difference solution on 129 x 129 points; its
    # own discretisation error reaches ~1e-2 in v near the right wall
    assert max(du) < 1e-2, max(du)
    assert max(dv) < 1e-2, max(dv)
    # the extrema of the converged solution (finer-grid and spectral computations after
    # Ghia et al. agree on u_min = -0.2140, v_max = 0.1796, v_min = -0.2538 at Re = 100;
    # taken here on the nodes, 1/128 apart, near the extremum positions)
    assert abs(min(uc.values()) - (-0.2140)) < 1.5e-3
    assert abs(max(vc.values()) - 0.1796) < 1.5e-3
    assert abs(min(vc.values()) - (-0.2538)) < 1.5e-3
