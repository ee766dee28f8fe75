"""EnforcedBC (Lagrange-multiplier) faces of the oracle, pinned by closed forms: the face
blocks are the 1D mass matrix, a Laplace problem whose exact solution is linear carries the
multiplier -du/dn exactly, and a constraint u = c gives the u of the Dirichlet problem
u = c; the problem Jacobian is the derivative of the problem residual."""
import numpy as np
import pytest

from oracle import assembly as oa
from oracle import faces as of
from oracle import meshes as om
from oracle import newton as on

CASES = [("C1", "left"), ("C2", "left"), ("C2", False)]


def _edge(nn, a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    pts = [a, b] if nn == 2 else [a, 0.5 * (a + b), b]
    return np.array([pts])


@pytest.mark.parametrize("nn,quad", [(2, False), (3, False), (3, True)])
def test_face_blocks_are_the_edge_mass_matrix(nn, quad):
    """J[u][l] = M and J[l][u] = alpha M with M the edge mass matrix; R[u] = M l and
    R[l] = alpha M u + beta * (edge load); u-u and l-l blocks zero."""
    X = _edge(nn, (0.3, -0.2), (1.1, 0.4))
    L = np.hypot(0.8, 0.6)
    rng = np.random.default_rng(1)
    U, Lm = rng.standard_normal((1, nn)), rng.standard_normal((1, nn))
    alpha, beta = 1.3, -0.7
    R, J = of.enforced_face_outputs(X, U, Lm, alpha, beta, 1, quad=quad)
    mass = np.array([[2, 1], [1, 2]]) / 6 if nn == 2 else np.array([[4, 2, -1], [2, 16, 2], [-1, 2, 4]]) / 30
    load = np.array([0.5, 0.5] if nn == 2 else [1 / 6, 2 / 3, 1 / 6])
    M = L * mass
    u_, l_ = slice(0, 2 * nn, 2), slice(1, 2 * nn, 2)
    assert np.allclose(J[0][u_, l_], M, rtol=1e-13, atol=1e-15)
    assert np.allclose(J[0][l_, u_], alpha * M, rtol=1e-13, atol=1e-15)
    assert np.all(J[0][u_, u_] == 0.0) and np.all(J[0][l_, l_] == 0.0)
    assert np.allclose(R[0][u_], M @ Lm[0], rtol=1e-13, atol=1e-14)
    assert np.allclose(R[0][l_], alpha * (M @ U[0]) + beta * L * load, rtol=1e-13, atol=1e-14)


def _problem(N, space, split, lam, dirichlet_left=1.0, enf=(1.0, -3.0), dirichlet_right=None):
    m = om.rectangular_quad_mesh(N, split, space)
    left, right = m.on_boundary(["left"]), m.on_boundary(["right"])
    if dirichlet_right is not None:
        pinned = (left | right)[:, None]
        bc = np.zeros((m.nnode, 1))
        bc[left, 0], bc[right, 0] = dirichlet_left, dirichlet_right
        return oa.OracleProblem(m, oa.FORM_BRATU, 1, pinned, bc, [lam]), m
    faces = of.boundary_faces(m, ["right"])
    on_f = np.zeros(m.nnode, bool)
    on_f[faces.reshape(-1)] = True
    pinned = np.zeros((m.nnode, 2), bool)
    pinned[:, 0] = left
    pinned[:, 1] = ~on_f | pinned[:, 0]
    bc = np.zeros((m.nnode, 2))
    bc[left, 0] = dirichlet_left
    op = oa.OracleProblem(m, oa.FORM_BRATU, 2, pinned, bc, [lam], enforced=faces, enf_alpha=enf[0],
                          enf_beta=enf[1])
    return op, m


@pytest.mark.parametrize("space,split", CASES)
def test_linear_solution_and_constant_multiplier(space, split):
    """-lap u = 0, u = 1 on the left, EnforcedBC(u = u - 3) on the right, natural elsewhere:
    u = 1 + 2x exactly, and the multiplier is -du/dn = -2 at every node of the right side
    (the boundary term of int grad u . grad v is int du/dn v, cancelled by int l v)."""
    op, m = _problem(6, space, split, 0.0)
    x = on.newton_solve(op, np.zeros(op.ndof))
    U = op.nodal_values(x)
    # Q2: the reference's Gauss<2,3> knot 0.774596662941483 (a typo of 0.774596669241483)
    # integrates the bulk stiffness inexactly, so linear fields are reproduced to ~1e-9 only
    tol = 1e-12 if split else 1e-8
    assert np.max(np.abs(U[:, 0] - (1.0 + 2.0 * m.coords[:, 0]))) < tol
    free_l = op.eqn[:, 1] >= 0
    assert free_l.sum() == m.on_boundary(["right"]).sum()
    assert np.max(np.abs(U[free_l, 1] + 2.0)) < 10 * tol


@pytest.mark.parametrize("space,split", CASES)
def test_constraint_reproduces_the_dirichlet_problem(space, split):
    """Bratu lambda = 1.2 with u = 0.3 on the right, enforced or pinned: the same u."""
    op, _ = _problem(5, space, split, 1.2, dirichlet_left=0.0, enf=(1.0, -0.3))
    od, _ = _problem(5, space, split, 1.2, dirichlet_left=0.0, dirichlet_right=0.3)
    ue = op.nodal_values(on.newton_solve(op, np.zeros(op.ndof)))[:, 0]
    ud = od.nodal_values(on.newton_solve(od, np.zeros(od.ndof)))[:, 0]
    assert np.max(np.abs(ue - ud)) < 1e-12


def test_problem_jacobian_is_the_residual_derivative():
    op, _ = _problem(4, "C2", "left", 1.5, enf=(1.7, -0.4))
    rng = np.random.default_rng(3)
    x = rng.uniform(-0.3, 0.3, op.ndof)
    _, rs, col, val = op.get_jacobian(x)
    import scipy.sparse as sp
    J = sp.csr_matrix((val, col, rs), shape=(op.ndof, op.ndof)).toarray()
    h = 1e-6
    for k in range(op.ndof):
        e = np.zeros(op.ndof)
        e[k] = h
        fd = (op.get_residuals(x + e) - op.get_residuals(x - e)) / (2 * h)
        assert np.max(np.abs(J[:, k] - fd)) < 1e-7 * max(1.0, np.max(np.abs(J[:, k])))
    # the multiplier rows and columns: no diagonal entry
    lrows = np.nonzero(op.eqn[:, 1] >= 0)[0]
    assert all(J[op.eqn[n, 1], op.eqn[n, 1]] == 0.0 for n in lrows)


@pytest.mark.parametrize("space,split", CASES)
def test_mesh_problem_numbering_matches_the_oracle(space, split):
    """Host numbering of the device problem: node by node, u then the multiplier, the
    multiplier only on the constrained side and pinned where u is pinned."""
    from pyoomph_amd.equations import BratuEquation
    from pyoomph_amd.meshes import RectangularQuadMesh
    from pyoomph_amd.problem import MeshProblem
    mp = MeshProblem(RectangularQuadMesh(5, split_in_tris=split, space=space),
                     BratuEquation("u", space=space, lambda_=1.0), dirichlet=[(("left", "top"), {0: 0.0})],
                     enforced=[(("right",), 1.0, -0.3)])
    m = om.rectangular_quad_mesh(5, split, space)
    faces = of.boundary_faces(m, ["right"])
    on_f = np.zeros(m.nnode, bool)
    on_f[faces.reshape(-1)] = True
    pinned = np.zeros((m.nnode, 2), bool)
    pinned[:, 0] = m.on_boundary(["left", "top"])
    pinned[:, 1] = ~on_f | pinned[:, 0]
    op = oa.OracleProblem(m, oa.FORM_BRATU, 2, pinned, np.zeros((m.nnode, 2)), [1.0], enforced=faces)
    assert mp.nvalue == 2 and np.array_equal(mp.eqn, op.eqn) and np.array_equal(mp.enforced, op.enforced)
    # the top-right corner: u pinned, so no multiplier
    corner = np.nonzero(m.on_boundary(["right"]) & m.on_boundary(["top"]))[0]
    assert np.all(mp.eqn[corner, 1] == -1)
    assert np.all(np.diff(mp.eqn[mp.eqn >= 0]) == 1)   # node-major order


def _expr_problem(N, space, split, lam, cexpr, C):
    """_problem with the constraint on the right side given by an expression."""
    op, m = _problem(N, space, split, lam)
    faces = of.boundary_faces(m, ["right"])
    op2 = oa.OracleProblem(m, oa.FORM_BRATU, 2, op.pinned, op.bc, [lam], enforced=faces,
                           enf_expr=cexpr, enf_coef=np.tile(np.asarray(C, float), (len(faces), 1)))
    return op2, m


@pytest.mark.parametrize("space,split", CASES)
def test_constraint_expressions(space, split):
    """Constraint expressions (the generated constraint forms): c0 u + c1 gives the built-in
    faces' residual and Jacobian exactly; u - x1^2 - 1 on the right of -lap u = 0 (u = 1 on
    the left) gives a solution equal to x1^2 + 1 on that side (the constraint holds pointwise
    up to the projection of the quadratic on the faces' space: exactly on P2 / Q2 faces); a
    nonlinear constraint with the parameter has the Jacobian and dR/dlambda of central
    differences."""
    lin, m = _expr_problem(4, space, split, 1.3, (lambda u, x0, x1, C: C[:, 0] * u + C[:, 1],
                                                  lambda u, x0, x1, C: C[:, 0] + 0.0 * u, None, None), (1.7, -0.4))
    ref, _ = _problem(4, space, split, 1.3, enf=(1.7, -0.4))
    x = np.random.default_rng(2).uniform(-0.3, 0.3, ref.ndof)
    r1, rs1, c1, v1 = lin.get_jacobian(x)
    r2, rs2, c2, v2 = ref.get_jacobian(x)
    assert np.array_equal(r1, r2) and np.array_equal(rs1, rs2) and np.array_equal(c1, c2) and np.array_equal(v1, v2)
    quad, m = _expr_problem(5, space, split, 0.0, (lambda u, x0, x1, C: u - x1 * x1 - C[:, 0],
                                                   lambda u, x0, x1, C: 1.0 + 0.0 * u, None, None), (1.0,))
    U = quad.nodal_values(on.newton_solve(quad, np.zeros(quad.ndof)))
    right = m.on_boundary(["right"])
    want = m.coords[right, 1] ** 2 + 1.0
    if space == "C2":
        assert np.max(np.abs(U[right, 0] - want)) < 1e-8
    else:
        assert np.max(np.abs(U[right, 0] - want)) < 0.05
    def cexpr(lam):
        return (lambda u, x0, x1, C: u ** 3 + C[:, 0] * x1 - lam * u * C[:, 1],
                lambda u, x0, x1, C: 3.0 * u ** 2 - lam * C[:, 1],
                lambda u, x0, x1, C: -u * C[:, 1],
                lambda u, x0, x1, C: -C[:, 1] + 0.0 * u)
    lam = 1.1
    nl, _ = _expr_problem(4, space, split, lam, cexpr(lam), (0.5, 2.0))
    x = np.random.default_rng(3).uniform(-0.4, 0.4, nl.ndof)
    _, rs, col, val = nl.get_jacobian(x)
    n = nl.ndof
    Jd = np.zeros((n, n))
    for i in range(n):
        for q in range(rs[i], rs[i + 1]):
            Jd[i, col[q]] += val[q]
    h = 1e-6
    for k in range(0, n, max(1, n // 17)):
        e = np.zeros(n)
        e[k] = h
        fd = (nl.get_residuals(x + e) - nl.get_residuals(x - e)) / (2 * h)
        assert np.allclose(Jd[:, k], fd, rtol=1e-7, atol=1e-9), k
    # dR/dlambda (the bulk's and the constraint's) against a central difference in lambda
    plus, _ = _expr_problem(4, space, split, lam + h, cexpr(lam + h), (0.5, 2.0))
    minus, _ = _expr_problem(4, space, split, lam - h, cexpr(lam - h), (0.5, 2.0))
    fd = (plus.get_residuals(x) - minus.get_residuals(x)) / (2 * h)
    assert np.allclose(nl.get_residuals(x, deriv=1), fd, rtol=1e-7, atol=1e-9)
