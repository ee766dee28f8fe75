"""Boundary-face flux elements of the oracle (NeumannBC(u=alpha*u + beta)), pinned by
closed forms: edge integrals, the 1D mass matrices, the Jacobian as the derivative of the
residual, and a Laplace problem with a Neumann / Robin side whose exact solution is
linear (reproduced exactly by P1 and P2)."""
import numpy as np
import pytest

from oracle import assembly as oa
from oracle import faces as of
from oracle import meshes as om
from oracle import newton as on


def _edge(nn, a, b):
    """one straight face from a to b (P2: midside node at the midpoint)"""
    a, b = np.asarray(a, float), np.asarray(b, float)
    pts = [a, b] if nn == 2 else [a, 0.5 * (a + b), b]
    return np.array([pts])


@pytest.mark.parametrize("nn,quad", [(2, False), (3, False), (3, True)])
def test_edge_load_and_mass_closed_forms(nn, quad):
    X = _edge(nn, (0.3, -0.2), (1.1, 0.4))
    L = np.hypot(0.8, 0.6)
    R, J = of.face_outputs(X, np.zeros((1, nn)), 0.0, 1.0, 1, quad=quad)
    load = [0.5, 0.5] if nn == 2 else [1 / 6, 2 / 3, 1 / 6]
    assert np.allclose(R[0], L * np.array(load), rtol=1e-13, atol=0)
    assert np.all(J == 0.0)
    _, M = of.face_outputs(X, np.zeros((1, nn)), 1.0, 0.0, 1, quad=quad)
    mass = np.array([[2, 1], [1, 2]]) / 6 if nn == 2 else np.array([[4, 2, -1], [2, 16, 2], [-1, 2, 4]]) / 30
    assert np.allclose(M[0], L * mass, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("nn", [2, 3])
def test_robin_residual_is_linear_and_jacobian_its_derivative(nn):
    rng = np.random.default_rng(5)
    X = _edge(nn, (0.0, 0.0), (0.25, 0.1))
    U = rng.standard_normal((1, nn))
    alpha, beta = 1.7, -0.4
    R, J = of.face_outputs(X, U, alpha, beta, 1)
    R0, _ = of.face_outputs(X, np.zeros((1, nn)), alpha, beta, 0)
    assert np.allclose(R[0], J[0] @ U[0] + R0[0], rtol=1e-13, atol=1e-15)
    h = 1e-6
    for j in range(nn):
        Up, Um = U.copy(), U.copy()
        Up[0, j] += h
        Um[0, j] -= h
        fd = (of.face_outputs(X, Up, alpha, beta, 0)[0] - of.face_outputs(X, Um, alpha, beta, 0)[0]) / (2 * h)
        assert np.allclose(fd[0], J[0][:, j], rtol=1e-7, atol=1e-10)


@pytest.mark.parametrize("space,split", [("C1", "left"), ("C2", "left"), ("C2", False)])
def test_boundary_faces_cover_each_side_once(space, split):
    N = 5
    m = om.rectangular_quad_mesh(N, split, space)
    for side in ("left", "right", "bottom", "top"):
        F = of.boundary_faces(m, [side])
        assert F.shape == (N, 2 if space == "C1" else 3)
        ends = m.coords[F[:, [0, -1]]]
        L = np.linalg.norm(ends[:, 1] - ends[:, 0], axis=1)
        assert np.isclose(L.sum(), 1.0, rtol=1e-14)
        on = m.on_boundary([side])
        assert np.all(on[F])


def test_product_faces_equal_oracle_faces():
    from pyoomph_amd import meshes as pm
    for space, split, kind in (("C1", "left", "tri_p1"), ("C2", "left", "tri_p2"), ("C2", False, "quad_q2")):
        m = om.rectangular_quad_mesh(6, split, space)
        pmesh = pm.RectangularQuadMesh(6, split, space)
        assert pmesh.kind == kind and np.array_equal(pmesh.elements, m.elements)
        for names in (["right"], ["top", "left"]):
            assert np.array_equal(pm.boundary_faces(pmesh, names), of.boundary_faces(m, names))


@pytest.mark.parametrize("space,split", [("C1", "left"), ("C2", "left"), ("C2", False)])
@pytest.mark.parametrize("alpha,beta", [(0.0, 0.7), (2.0, 0.7)])
def test_laplace_with_flux_side_has_linear_solution(space, split, alpha, beta):
    """-lap u = 0 (Bratu with lambda = 0), u = 0 on the left, flux alpha*u + beta on the
    right (NeumannBC: du/dn = -(alpha*u + beta)), natural top and bottom: u = c x with
    c = -beta / (1 + alpha), which P1, P2 and Q2 reproduce."""
    m = om.rectangular_quad_mesh(4, split, space)
    pinned = m.on_boundary({"left"})[:, None]
    F = of.boundary_faces(m, ["right"])
    op = oa.OracleProblem(m, oa.FORM_BRATU, 1, pinned, np.zeros((m.nnode, 1)), [0.0], faces=F,
                          face_alpha=alpha, face_beta=beta)
    x = on.newton_solve(op, np.zeros(op.ndof), tol=1e-12)
    u = op.nodal_values(x)[:, 0]
    c = -beta / (1.0 + alpha)
    # Q2 bulk elements integrate with the reference's Gauss<2,3>, whose positive knot is
    # mistyped (0.774596662941483): the linear field is reproduced to that knot's ~1e-8
    tol = 1e-12 if split else 5e-9
    assert np.allclose(u, c * m.coords[:, 0], rtol=0, atol=tol)


def test_face_contributions_follow_the_bulk_in_the_csr():
    """Face couplings lie inside their bulk element's couplings: the pattern is the bulk
    pattern, and the values gain the face integrals (Robin: the face mass matrix)."""
    m = om.rectangular_quad_mesh(3, "left", "C2")
    pinned = m.on_boundary({"left"})[:, None]
    F = of.boundary_faces(m, ["right", "top"])
    kw = dict(mesh=m, form=oa.FORM_BRATU, nvalue=1, pinned=pinned, bc=np.zeros((m.nnode, 1)), params=[1.0])
    bulk = oa.OracleProblem(**kw)
    face = oa.OracleProblem(**kw, faces=F, face_alpha=3.0, face_beta=0.0)
    x = np.random.default_rng(2).standard_normal(bulk.ndof) * 0.1
    rb, rsb, cb, vb = bulk.get_jacobian(x)
    rf, rsf, cf, vf = face.get_jacobian(x)
    assert np.array_equal(rsb, rsf) and np.array_equal(cb, cf)
    # the difference is alpha times the assembled face mass matrix, and R changes by J_face x
    import scipy.sparse as sp
    n = bulk.ndof
    Jb = sp.csr_matrix((vb, cb, rsb), shape=(n, n)).toarray()
    Jf = sp.csr_matrix((vf, cf, rsf), shape=(n, n)).toarray()
    D = Jf - Jb
    assert np.allclose(D, D.T, atol=1e-15) and np.all(np.diag(D) >= 0) and D.sum() > 0
    assert np.allclose(rf - rb, D @ x, rtol=1e-12, atol=1e-15)


def test_several_fluxes_on_one_mesh_superpose():
    """Per-face coefficients: Neumann on one side and Robin on another give the sum of
    the two conditions' face terms (residual and Jacobian)."""
    m = om.rectangular_quad_mesh(4, "left", "C1")
    pinned = m.on_boundary({"left"})[:, None]
    kw = dict(mesh=m, form=oa.FORM_BRATU, nvalue=1, pinned=pinned, bc=np.zeros((m.nnode, 1)), params=[0.8])
    Fr, Ft = of.boundary_faces(m, ["right"]), of.boundary_faces(m, ["top"])
    both = oa.OracleProblem(**kw, faces=np.concatenate([Fr, Ft]),
                            face_alpha=np.r_[np.zeros(len(Fr)), np.full(len(Ft), 2.5)],
                            face_beta=np.r_[np.full(len(Fr), 0.4), np.full(len(Ft), -0.3)])
    right = oa.OracleProblem(**kw, faces=Fr, face_alpha=0.0, face_beta=0.4)
    top = oa.OracleProblem(**kw, faces=Ft, face_alpha=2.5, face_beta=-0.3)
    none = oa.OracleProblem(**kw)
    x = np.random.default_rng(4).standard_normal(none.ndof) * 0.2
    import scipy.sparse as sp
    n = none.ndof

    def dense(op):
        r, rs, c, v = op.get_jacobian(x)
        return r, sp.csr_matrix((v, c, rs), shape=(n, n)).toarray()
    (rb, Jb), (rr, Jr), (rt, Jt), (r0, J0) = dense(both), dense(right), dense(top), dense(none)
    assert np.allclose(rb - r0, (rr - r0) + (rt - r0), rtol=1e-12, atol=1e-14)
    assert np.allclose(Jb - J0, (Jr - J0) + (Jt - J0), rtol=1e-12, atol=1e-14)
    assert np.array_equal(Jr, J0)            # a Neumann flux adds nothing to the Jacobian


# The reference's Gauss<2,3> (the rule of Q2 quadrilaterals and of the QElement<2,3> faces of
# Q2 bricks) has a mistyped positive knot (0.774596662941483): its integrals of the exact
# polynomials are off by up to ~1e-7 relative (entry by entry), which bounds these checks.
Q9_TOL = 2e-7


def test_brick_face_load_and_mass_closed_forms():
    """a planar square face (side h, tilted out of the coordinate planes): load
    h^2 (1/6, 2/3, 1/6) x (1/6, 2/3, 1/6), mass h^2 M1 x M1, M1 = [[4,2,-1],[2,16,2],[-1,2,4]]/30"""
    h, th = 0.4, 0.3
    X = np.array([[[0.1 + a * h / 2, 0.2 + b * h / 2 * np.cos(th), b * h / 2 * np.sin(th)]
                   for b in range(3) for a in range(3)]])
    v = np.array([1 / 6, 2 / 3, 1 / 6])
    R, J = of.face_outputs_quad9(X, np.zeros((1, 9)), 0.0, 1.0, 1)
    assert np.allclose(R[0], h * h * np.kron(v, v), rtol=Q9_TOL, atol=0) and np.all(J == 0.0)
    _, M = of.face_outputs_quad9(X, np.zeros((1, 9)), 1.0, 0.0, 1)
    M1 = np.array([[4, 2, -1], [2, 16, 2], [-1, 2, 4]]) / 30
    assert np.allclose(M[0], h * h * np.kron(M1, M1), rtol=0, atol=Q9_TOL * h * h)
    U = np.random.default_rng(8).standard_normal((1, 9))
    Rr, Jr = of.face_outputs_quad9(X, U, 1.3, -0.2, 1)
    R0, _ = of.face_outputs_quad9(X, np.zeros((1, 9)), 1.3, -0.2, 0)
    assert np.allclose(Rr[0], Jr[0] @ U[0] + R0[0], rtol=1e-13, atol=1e-15)


def test_brick_boundary_faces():
    from pyoomph_amd import meshes as pm
    N = 3
    m = om.cuboid_brick_mesh(N)
    for side in ("left", "right", "bottom", "top", "back", "front"):
        F = of.boundary_faces(m, [side])
        assert F.shape == (N * N, 9) and np.all(m.on_boundary([side])[F])
    assert np.array_equal(pm.boundary_faces(pm.CuboidBrickMesh(N), ["left", "front"]),
                          of.boundary_faces(m, ["left", "front"]))


@pytest.mark.parametrize("alpha,beta", [(0.0, 0.7), (2.0, 0.7)])
def test_laplace_in_a_cube_with_flux_side_has_linear_solution(alpha, beta):
    m = om.cuboid_brick_mesh(2)
    pinned = m.on_boundary({"left"})[:, None]
    op = oa.OracleProblem(m, oa.FORM_BRATU, 1, pinned, np.zeros((m.nnode, 1)), [0.0],
                          faces=of.boundary_faces(m, ["right"]), face_alpha=alpha, face_beta=beta)
    x = on.newton_solve(op, np.zeros(op.ndof), tol=1e-12)
    u = op.nodal_values(x)[:, 0]
    assert np.allclose(u, -beta / (1.0 + alpha) * m.coords[:, 0], rtol=0, atol=1e-8)


def _ns_traction_problem(N, alpha, beta, faces=True):
    """Taylor-Hood on the unit square: u = (1, 0) on the top, u = 0 on the left and the
    bottom, the traction alpha*u + beta on the right (NeumannBC on the velocity components),
    the pressure pinned at node 0 (a one-sided outflow would fix its level anyway)."""
    m = om.rectangular_quad_mesh(N, False, "C2")
    nn = m.nnode
    vert = np.zeros(nn, bool)
    vert[m.elements[:, [0, 2, 6, 8]].ravel()] = True
    top, walls = m.on_boundary({"top"}), m.on_boundary({"left", "bottom"})
    pinned = np.zeros((nn, 3), bool)
    bc = np.zeros((nn, 3))
    pinned[top | walls, :2] = True
    bc[top & ~walls, 0] = 1.0
    pinned[~vert, 2] = True
    pinned[0, 2] = True
    F = of.boundary_faces(m, ["right"]) if faces else None
    return m, oa.OracleProblem(m, oa.FORM_NS, 3, pinned, bc, [0.05, 1.0], faces=F, face_alpha=alpha,
                               face_beta=np.asarray(beta, float))


def test_taylor_hood_traction_load_closed_form():
    """alpha = 0: the residual gains the traction load; summed over the free x- (y-) momentum
    rows it is beta_x (beta_y) times the right side's length less the two corner nodes'
    shares (h/6 each: the top and bottom corners are Dirichlet nodes)."""
    N = 4
    beta = (0.3, -0.2)
    m, op = _ns_traction_problem(N, 0.0, beta)
    _, op0 = _ns_traction_problem(N, 0.0, beta, faces=False)
    x = np.random.default_rng(1).standard_normal(op.ndof) * 0.1
    d = op.get_residuals(x) - op0.get_residuals(x)
    for c in range(2):
        rows = op.eqn[:, c][op.eqn[:, c] >= 0]
        assert np.isclose(d[rows].sum(), beta[c] * (1.0 - 2 * (1.0 / N) / 6), rtol=1e-12)
    assert np.allclose(d[op.eqn[:, 2][op.eqn[:, 2] >= 0]], 0.0)     # no pressure rows
    # alpha != 0: the Jacobian gains alpha times the face mass matrix on each component
    _, opa = _ns_traction_problem(N, 2.0, beta)
    _, rs, col, val = opa.get_jacobian(x)
    _, rs0, col0, val0 = op.get_jacobian(x)
    import scipy.sparse as sp
    n = op.ndof
    D = sp.csr_matrix((val, col, rs), shape=(n, n)).toarray() - sp.csr_matrix((val0, col0, rs0), shape=(n, n)).toarray()
    ux, uy = op.eqn[:, 0], op.eqn[:, 1]
    assert np.allclose(D, D.T, atol=1e-15)
    assert np.all(D[np.ix_(ux[ux >= 0], uy[uy >= 0])] == 0.0)       # components do not couple


@pytest.mark.parametrize("nn,quad", [(2, False), (3, False), (3, True)])
def test_expression_faces_closed_form_and_derivative(nn, quad):
    """Expression fluxes (the generated boundary-face forms): a flux x0 on the edge from
    (0, 0) to (L, 0) integrates to the closed-form loads int x psi_i ds; the Jacobian of a
    nonlinear flux is the central difference of the residual."""
    L = 1.3
    X = _edge(nn, (0.0, 0.0), (L, 0.0))
    C = np.zeros((1, 0))
    R, _ = of.expression_face_outputs(X, np.zeros((1, nn)), C, lambda u, x0, x1, C: x0,
                                      lambda u, x0, x1, C: 0.0 * u, 1, quad=quad)
    exact = [L * L / 6, L * L / 3] if nn == 2 else [0.0, L * L / 3, L * L / 6]
    assert np.allclose(R[0], exact, rtol=1e-13, atol=1e-15)
    rng = np.random.default_rng(4)
    X = _edge(nn, (0.2, 0.1), (0.9, 0.5))
    C = np.array([[1.7, -0.3]])
    F = lambda u, x0, x1, C: C[:, 0] * (u * u * u) + C[:, 1] * x1 * u  # noqa: E731
    D = lambda u, x0, x1, C: C[:, 0] * (3.0 * (u * u)) + C[:, 1] * x1  # noqa: E731
    U = rng.uniform(-0.5, 0.5, (1, nn))
    R, J = of.expression_face_outputs(X, U, C, F, D, 1, quad=quad)
    h = 1e-6
    for k in range(nn):
        e = np.zeros((1, nn))
        e[0, k] = h
        Rp, _ = of.expression_face_outputs(X, U + e, C, F, D, 0, quad=quad)
        Rm, _ = of.expression_face_outputs(X, U - e, C, F, D, 0, quad=quad)
        assert np.allclose(J[0][:, k], (Rp[0] - Rm[0]) / (2 * h), rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("space,split", [("C1", "left"), ("C2", "left"), ("C2", False)])
def test_outward_normal_of_the_unit_square(space, split):
    """The normal the expression faces get: on each side of the unit square, the known
    outward normal (+-1, 0) / (0, +-1) at every point, whatever the face's node order."""
    from oracle import meshes as om
    m = om.rectangular_quad_mesh(4, split, space)
    want = {"right": (1.0, 0.0), "left": (-1.0, 0.0), "top": (0.0, 1.0), "bottom": (0.0, -1.0)}
    for name, (w0, w1) in want.items():
        F = of.boundary_faces(m, [name])
        sides = of.outward_sides(m.coords, m.elements, F)
        X = np.ascontiguousarray(m.coords[F][:, :, :2])
        got = []
        of.expression_face_outputs(X, np.zeros((len(F), F.shape[1])), np.zeros((len(F), 0)),
                                   lambda u, x0, x1, C, n0, n1: got.append((n0, n1)) or 0.0 * u,
                                   lambda u, x0, x1, C, n0, n1: 0.0 * u, 0, quad=m.kind == "quad_q2",
                                   sides=sides)
        for n0, n1 in got:
            assert np.allclose(n0, w0, atol=1e-15) and np.allclose(n1, w1, atol=1e-15), name


def test_normal_flux_equals_the_per_side_constant_flux():
    """NeumannBC(u = c . n) on all four sides: the same residual as constant fluxes c0, -c0,
    c1, -c1 on the right, left, top and bottom faces."""
    from oracle import meshes as om
    m = om.rectangular_quad_mesh(5, "left", "C2")
    c = np.array([0.7, -1.3])
    names = ["right", "left", "top", "bottom"]
    faces = of.boundary_faces(m, names)
    C = np.tile(c, (len(faces), 1))
    pinned = np.zeros((m.nnode, 1), bool)
    pinned[0, 0] = True
    kw = dict(faces=faces, face_coef=C)
    normal = oa.OracleProblem(m, oa.FORM_POISSON, 1, pinned, np.zeros((m.nnode, 1)), [0],
                              face_expr=(lambda u, x0, x1, C, n0, n1: C[:, 0] * n0 + C[:, 1] * n1,
                                         lambda u, x0, x1, C, n0, n1: 0.0 * u, None, None),
                              face_normal=True, **kw)
    per_side = np.concatenate([np.full(len(of.boundary_faces(m, [n])), v)
                               for n, v in zip(names, (c[0], -c[0], c[1], -c[1]))])
    const = oa.OracleProblem(m, oa.FORM_POISSON, 1, pinned, np.zeros((m.nnode, 1)), [0],
                             face_expr=(lambda u, x0, x1, C: C[:, 0], lambda u, x0, x1, C: 0.0 * u, None, None),
                             faces=faces, face_coef=per_side[:, None])
    x = np.random.default_rng(1).uniform(-1, 1, normal.ndof)
    assert np.allclose(normal.get_residuals(x), const.get_residuals(x), rtol=0, atol=1e-14)
    assert np.max(np.abs(normal.get_residuals(x) - oa.OracleProblem(
        m, oa.FORM_POISSON, 1, pinned, np.zeros((m.nnode, 1)), [0]).get_residuals(x))) > 1e-3
