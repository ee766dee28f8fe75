"""Pins the oracle's element functions and CSR builder.

* stiffness matrices against closed forms: P1 (constant gradients), P2 (barycentric
  integration formula), Q2 quads and Q2 bricks (1D stiffness/mass tensor products);
* Jacobians against central finite differences of the residual (the invariant of
  pyoomph's debug_analytical_jacobian, src/elements.cpp:5880) for every nonlinear form;
* dR/dlambda against finite differences in lambda;
* the vectors_of_pairs CSR rules (first-touch column order, exact zeros dropped per
  contribution, element-order sums) on hand-built element matrices;
* brute-force dense assembly.
"""
import math
from itertools import product

import numpy as np
import pytest

from oracle import assembly as oa
from oracle import meshes as om
from tests.oracle_bridge import oracle_problem


def element_matrix(op, e, x=None):
    x = np.zeros(op.ndof) if x is None else x
    R, J = op.element_outputs(x, 1)
    nv = op.nvar[e]
    return J[e].reshape(-1)[:nv * nv].reshape(nv, nv)


def free_problem(mesh, form, params=(), aux=None):
    """All nodes free (no Dirichlet): element local order == node order."""
    pinned = np.zeros((mesh.nnode, 1), bool)
    return oa.OracleProblem(mesh, form, 1, pinned, np.zeros((mesh.nnode, 1)), params, aux=aux)


def p2_stiffness_exact(X):
    """Closed form: grad phi_vertex_a = (4 l_a - 1) grad l_a, grad phi_edge_ab =
    4 (l_a grad l_b + l_b grad l_a); integrals of barycentric monomials
    int l1^a l2^b l3^c = 2A a! b! c! / (a+b+c+2)!."""
    X = np.asarray(X, float)
    T = np.array([[X[0, 0] - X[2, 0], X[1, 0] - X[2, 0]], [X[0, 1] - X[2, 1], X[1, 1] - X[2, 1]]])
    A = abs(np.linalg.det(T)) / 2
    G = np.linalg.inv(T).T            # columns: grad l0, grad l1 ; l2 = 1 - l0 - l1
    g = [G[:, 0], G[:, 1], -G[:, 0] - G[:, 1]]

    def integ(powers):
        a = list(powers)
        return 2 * A * math.factorial(a[0]) * math.factorial(a[1]) * math.factorial(a[2]) / \
            math.factorial(sum(a) + 2)

    # each gradient is a sum of terms coef * l_k (power vector) * vector
    def grad_terms(i):
        if i < 3:
            e = [0, 0, 0]
            e[i] = 1
            return [(4.0, tuple(e), g[i]), (-1.0, (0, 0, 0), g[i])]
        a, b = {3: (0, 1), 4: (1, 2), 5: (2, 0)}[i]
        ea = [0, 0, 0]
        ea[a] = 1
        eb = [0, 0, 0]
        eb[b] = 1
        return [(4.0, tuple(ea), g[b]), (4.0, tuple(eb), g[a])]

    K = np.zeros((6, 6))
    for i in range(6):
        for j in range(6):
            s = 0.0
            for ci, pi, vi in grad_terms(i):
                for cj, pj, vj in grad_terms(j):
                    s += ci * cj * (vi @ vj) * integ(np.add(pi, pj))
            K[i, j] = s
    return K


def test_p1_stiffness_closed_form():
    m = om.rectangular_quad_mesh(1, "left", "C1")
    op = free_problem(m, oa.FORM_POISSON, [0.5, 0.5, 1.0, 0.0])
    for e in range(m.nelement):
        X = m.coords[m.elements[e]]
        T = np.array([[X[0, 0] - X[2, 0], X[1, 0] - X[2, 0]], [X[0, 1] - X[2, 1], X[1, 1] - X[2, 1]]])
        G = np.linalg.inv(T).T
        g = np.stack([G[:, 0], G[:, 1], -G[:, 0] - G[:, 1]])
        K = abs(np.linalg.det(T)) / 2 * g @ g.T
        assert np.allclose(element_matrix(op, e), K, atol=1e-12)


def test_p2_stiffness_closed_form():
    m = om.rectangular_quad_mesh(2, "left", "C2")
    m.coords[:, 0] *= 1.7   # non-square aspect
    op = free_problem(m, oa.FORM_POISSON, [0.5, 0.5, 1.0, 0.0])
    for e in range(m.nelement):
        K = p2_stiffness_exact(m.coords[m.elements[e]][:3])
        assert np.allclose(element_matrix(op, e), K, atol=1e-11)


def _lag_1d(h):
    # 1D quadratic Lagrange stiffness / mass on [0, h], nodes 0, h/2, h
    K = np.array([[7, -8, 1], [-8, 16, -8], [1, -8, 7]]) / (3 * h)
    M = np.array([[4, 2, -1], [2, 16, 2], [-1, 2, 4]]) * h / 30
    return K, M


def test_q2_stiffness_tensor_product():
    m = om.rectangular_quad_mesh(1, False, "C2")
    m.coords[:, 1] *= 0.5
    op = free_problem(m, oa.FORM_POISSON, [0.5, 0.5, 1.0, 0.0])
    Kx, Mx = _lag_1d(1.0)
    Ky, My = _lag_1d(0.5)
    K = np.kron(My, Kx) + np.kron(Ky, Mx)   # node index = iy*3 + ix
    # Gauss<2,3>'s mistyped knot makes the rule inexact by ~1e-8 relative
    assert np.allclose(element_matrix(op, 0), K, rtol=0, atol=5e-8)


def test_hex_stiffness_tensor_product():
    m = om.cuboid_brick_mesh(1)
    op = free_problem(m, oa.FORM_RD, [], aux=np.zeros(m.nnode))
    K1, M1 = _lag_1d(1.0)
    K = (np.kron(np.kron(M1, M1), K1) + np.kron(np.kron(M1, K1), M1) + np.kron(np.kron(K1, M1), M1))
    assert np.allclose(element_matrix(op, 0), K, atol=1e-12)


def _fd_jacobian(op, x, h=1e-6):
    n = op.ndof
    J = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = h
        J[:, j] = (op.get_residuals(x + e) - op.get_residuals(x - e)) / (2 * h)
    return J


@pytest.mark.parametrize("kind,kw", [("bratu", dict(N=3, space="C1", lam=2.0)),
                                     ("bratu", dict(N=2, space="C2", lam=2.0)),
                                     ("poisson", dict(N=2)),
                                     ("ns", dict(N=2))])
def test_jacobian_matches_finite_differences(kind, kw):
    op = oracle_problem(kind, **kw)
    x = np.random.default_rng(2).uniform(-0.5, 0.5, op.ndof)
    r, rs, col, val = op.get_jacobian(x)
    J = np.zeros((op.ndof, op.ndof))
    for i in range(op.ndof):
        J[i, col[rs[i]:rs[i + 1]]] = val[rs[i]:rs[i + 1]]
    Jfd = _fd_jacobian(op, x)
    assert np.max(np.abs(J - Jfd)) < 1e-7 * max(1.0, np.max(np.abs(J)))
    assert np.allclose(r, op.get_residuals(x), atol=0, rtol=0)


def test_rd_jacobian_matches_finite_differences():
    m = om.cuboid_brick_mesh(1)
    f = np.linspace(-1, 1, m.nnode)
    op = free_problem(m, oa.FORM_RD, [], aux=f)
    x = np.random.default_rng(5).uniform(-1, 1, op.ndof)
    r, rs, col, val = op.get_jacobian(x)
    J = np.zeros((op.ndof, op.ndof))
    for i in range(op.ndof):
        J[i, col[rs[i]:rs[i + 1]]] = val[rs[i]:rs[i + 1]]
    assert np.max(np.abs(J - _fd_jacobian(op, x))) < 1e-7


def test_bratu_lambda_derivative():
    op = oracle_problem("bratu", N=3, space="C2", lam=1.3)
    x = np.random.default_rng(3).uniform(0, 0.3, op.ndof)
    d = op.get_residuals(x, deriv=1)
    h = 1e-6
    op.params[0] = 1.3 + h
    rp = op.get_residuals(x)
    op.params[0] = 1.3 - h
    rm = op.get_residuals(x)
    op.params[0] = 1.3
    assert np.allclose(d, (rp - rm) / (2 * h), atol=1e-9)


def test_dense_brute_force_assembly():
    op = oracle_problem("bratu", N=2, space="C2", lam=1.0)
    x = np.random.default_rng(4).uniform(-0.2, 0.2, op.ndof)
    R, J = op.element_outputs(x, 1)
    D = np.zeros((op.ndof, op.ndof))
    rr = np.zeros(op.ndof)
    for e in range(op.mesh.nelement):
        nv = op.nvar[e]
        g = op.geqn[e, :nv]
        Je = J[e].reshape(-1)[:nv * nv].reshape(nv, nv)
        D[np.ix_(g, g)] += Je
        rr[g] += R[e, :nv]
    r, rs, col, val = op.get_jacobian(x)
    S = np.zeros_like(D)
    for i in range(op.ndof):
        S[i, col[rs[i]:rs[i + 1]]] = val[rs[i]:rs[i + 1]]
    assert np.allclose(S, D, atol=1e-15)
    assert np.allclose(r, rr, atol=1e-15)


def test_vectors_of_pairs_rules():
    """Hand-built element matrices: first-touch column order, exact zeros dropped per
    contribution (even where a later contribution makes the entry non-zero, its column
    moves to the time of that first non-zero touch), element-order sums."""
    import ctypes
    L = oa.lib()
    nel, maxdof, ndof = 2, 3, 4
    nvar = np.array([3, 3], dtype=np.int64)
    geqn = np.array([[2, 0, 1], [1, 3, 2]], dtype=np.int64)
    J = np.zeros((nel, maxdof, maxdof))
    J[0] = [[1.0, 0.0, 2.0],       # row eqn 2: cols 2 (1.0), 0 (0.0 dropped), 1 (2.0)
            [3.0, 4.0, 0.0],       # row eqn 0: cols 2, 0 ; col 1 dropped
            [0.0, 5.0, 6.0]]       # row eqn 1: col 0, 1
    J[1] = [[7.0, 8.0, 0.0],       # row eqn 1: col 1 (+=7), col 3 new, col 2 dropped
            [0.0, 9.0, 10.0],      # row eqn 3: col 3, col 2
            [11.0, 0.0, -1.0]]     # row eqn 2: col 1 (+=11), col 2 (+= -1 -> 0.0 kept)
    R = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]])
    res = np.zeros(ndof)
    rs = np.zeros(ndof + 1, dtype=np.int64)
    colp = ctypes.POINTER(ctypes.c_long)()
    valp = ctypes.POINTER(ctypes.c_double)()
    nnz = L.oracle_vectors_of_pairs(ndof, nel, maxdof, oa._lp(nvar), oa._lp(geqn), oa._dp(R),
                                    oa._dp(np.ascontiguousarray(J)), oa._dp(res), oa._lp(rs),
                                    ctypes.byref(colp), ctypes.byref(valp))
    col = np.ctypeslib.as_array(colp, shape=(nnz,)).copy()
    val = np.ctypeslib.as_array(valp, shape=(nnz,)).copy()
    rows = {i: list(zip(col[rs[i]:rs[i + 1]], val[rs[i]:rs[i + 1]])) for i in range(ndof)}
    assert rows[0] == [(2, 3.0), (0, 4.0)]
    assert rows[1] == [(0, 5.0), (1, 6.0 + 7.0), (3, 8.0)]
    assert rows[2] == [(2, 1.0 - 1.0), (1, 2.0 + 11.0)]
    assert rows[3] == [(3, 9.0), (2, 10.0)]
    assert res.tolist() == [2.0, 3.0 + 4.0, 1.0 + 6.0, 5.0]


@pytest.mark.parametrize("kind,kw", [("bratu", dict(N=3, space="C1", lam=2.0)),
                                     ("bratu", dict(N=2, space="C2", lam=1.3)),
                                     ("poisson", dict(N=2))])
def test_hessian_vector_product_matches_finite_differences(kind, kw):
    """sum_k dJ/du_k y_k against (J(x + h y) - J(x - h y)) / 2h."""
    op = oracle_problem(kind, **kw)
    rng = np.random.default_rng(8)
    x, y = rng.uniform(-0.5, 0.5, op.ndof), rng.uniform(-1, 1, op.ndof)

    def dense_j(v, **kw2):
        _, rs, col, val = op.get_jacobian(v, **kw2)
        M = np.zeros((op.ndof, op.ndof))
        for i in range(op.ndof):
            M[i, col[rs[i]:rs[i + 1]]] = val[rs[i]:rs[i + 1]]
        return M

    H = dense_j(x, y=y)
    h = 1e-5
    Hfd = (dense_j(x + h * y) - dense_j(x - h * y)) / (2 * h)
    assert np.max(np.abs(H - Hfd)) < 1e-7 * max(1.0, np.max(np.abs(H)))
    if kind == "poisson":
        assert not np.any(H)


def test_rd_hessian_vector_product_matches_finite_differences():
    m = om.cuboid_brick_mesh(1)
    op = free_problem(m, oa.FORM_RD, [], aux=np.linspace(-1, 1, m.nnode))
    rng = np.random.default_rng(9)
    x, y = rng.uniform(-1, 1, op.ndof), rng.uniform(-1, 1, op.ndof)
    Jp = op.element_outputs(x + 1e-5 * y, 1)[1][0]
    Jm = op.element_outputs(x - 1e-5 * y, 1)[1][0]
    H = op.element_outputs(x, 1, y=y)[1][0]
    assert np.max(np.abs(H - (Jp - Jm) / 2e-5)) < 1e-7 * max(1.0, np.max(np.abs(H)))


def test_taylor_hood_hessian_vector_product_matches_finite_differences():
    """Taylor-Hood: the Jacobian is linear in the velocity (the convection term is the only
    non-linearity), so (J(x + h y) - J(x - h y)) / 2h equals sum_k dJ/du_k y_k up to
    rounding; pressure rows and columns of the product vanish."""
    op = oracle_problem("ns", N=2, Re=50.0)
    rng = np.random.default_rng(12)
    x, y = rng.uniform(-0.5, 0.5, op.ndof), rng.uniform(-1, 1, op.ndof)

    def dense_j(v, **kw2):
        _, rs, col, val = op.get_jacobian(v, **kw2)
        M = np.zeros((op.ndof, op.ndof))
        for i in range(op.ndof):
            M[i, col[rs[i]:rs[i + 1]]] = val[rs[i]:rs[i + 1]]
        return M

    H = dense_j(x, y=y)
    h = 1e-3
    Hfd = (dense_j(x + h * y) - dense_j(x - h * y)) / (2 * h)
    assert np.max(np.abs(H)) > 1e-3
    assert np.max(np.abs(H - Hfd)) < 1e-9 * max(1.0, np.max(np.abs(H)))
    p = op.eqn[:, 2][op.eqn[:, 2] >= 0]
    assert not H[p].any() and not H[:, p].any()
