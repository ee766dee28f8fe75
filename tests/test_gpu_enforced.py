"""GPU EnforcedBC (Lagrange-multiplier) faces against the oracle: the reference CSR of the
saddle-point Jacobian (pattern bit-exact; values bitwise without the Bratu exponential,
else within 1e-12), Newton with GMRES + ILU(0) (multipliers ordered last) against the
oracle's SuperLU Newton and against the closed form of a linear solution, and a batch
equal to each sub-problem alone."""
import numpy as np
import pytest

from oracle import assembly as oa
from oracle import faces as of
from oracle import meshes as om
from oracle import newton as on

pytestmark = pytest.mark.gpu

CASES = [("C1", "left"), ("C2", "left"), ("C2", False)]


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def build(N, space, split, lam, enf=(1.0, -3.0), left=1.0, flux=None):
    from pyoomph_amd.equations import BratuEquation
    from pyoomph_amd.meshes import RectangularQuadMesh
    from pyoomph_amd.problem import MeshProblem
    mp = MeshProblem(RectangularQuadMesh(N, split_in_tris=split, space=space),
                     BratuEquation("u", space=space, lambda_=lam), dirichlet=[(("left",), {0: left})],
                     enforced=[(("right",), enf[0], enf[1])], flux=flux)
    m = om.rectangular_quad_mesh(N, split, space)
    faces = of.boundary_faces(m, ["right"])
    on_f = np.zeros(m.nnode, bool)
    on_f[faces.reshape(-1)] = True
    pinned = np.zeros((m.nnode, 2), bool)
    pinned[:, 0] = m.on_boundary(["left"])
    pinned[:, 1] = ~on_f | pinned[:, 0]
    bc = np.zeros((m.nnode, 2))
    bc[pinned[:, 0], 0] = left
    kw = {}
    if flux is not None:
        (names, a, b), = flux
        kw = dict(faces=of.boundary_faces(m, list(names)), face_alpha=a, face_beta=b)
    op = oa.OracleProblem(m, oa.FORM_BRATU, 2, pinned, bc, [lam], enforced=faces, enf_alpha=enf[0],
                          enf_beta=enf[1], **kw)
    assert np.array_equal(mp.eqn, op.eqn) and np.array_equal(mp.enforced, op.enforced)
    return mp, op


def solver(P):
    P.krylov, P.preconditioner = "gmres", "ilu0"
    P.gmres_restart, P.gmres_rtol = 100, 1e-12
    P.newton_solver_tolerance = 1e-11
    return P


@pytest.mark.parametrize("space,split", CASES)
@pytest.mark.parametrize("lam", [0.0, 1.5])
def test_assembly_parity(gpu, space, split, lam):
    from pyoomph_amd.problem import Problem
    mp, op = build(6, space, split, lam, enf=(1.7, -0.4))
    P = Problem(mp)
    for seed in (1, 2):
        x = np.random.default_rng(seed).uniform(-0.4, 0.4, P.ndof)
        P.set_current_dofs(x)
        R, J = P.assemble_jacobian()
        r, rs, col, val = op.get_jacobian(x)
        assert np.array_equal(J.indptr, rs) and np.array_equal(J.indices, col)
        if lam == 0.0:
            assert np.array_equal(J.data, val) and np.array_equal(R, r)
        else:
            assert rel(J.data, val) <= 1e-12 and rel(R, r) <= 1e-12
    # no multiplier row keeps a diagonal entry: exact zeros are dropped as the reference drops them
    lrow = mp.eqn[mp.eqn[:, 1] >= 0, 1]
    assert not np.any(J.toarray()[lrow, lrow])


@pytest.mark.parametrize("space,split", CASES)
def test_linear_solution_and_multiplier(gpu, space, split):
    """u = 1 on the left, EnforcedBC(u = u - 3) on the right, -lap u = 0: u = 1 + 2x and the
    multiplier -du/dn = -2 on the right side (Q2: to the accuracy of the Gauss<2,3> knot)."""
    from pyoomph_amd.problem import Problem
    mp, op = build(8, space, split, 0.0)
    P = solver(Problem(mp))
    st = P.solve()
    assert st[0]["converged"]
    U = mp.nodal_values(P.get_current_dofs())
    tol = 1e-10 if split else 1e-8
    assert np.max(np.abs(U[:, 0] - (1.0 + 2.0 * mp.mesh.coords[:, 0]))) < tol
    assert np.max(np.abs(U[mp.eqn[:, 1] >= 0, 1] + 2.0)) < 10 * tol


@pytest.mark.parametrize("space,split", CASES)
def test_newton_matches_oracle(gpu, space, split):
    from pyoomph_amd.problem import Problem
    mp, op = build(10, space, split, 1.2, enf=(1.0, -0.3), left=0.0)
    P = solver(Problem(mp))
    st = P.solve()
    assert st[0]["converged"] and st[0]["linear_failures"] == 0
    xo = on.newton_solve(op, np.zeros(op.ndof), tol=1e-12)
    assert rel(P.get_current_dofs(), xo) <= 1e-8


def test_with_flux_faces_and_batch_equals_alone(gpu):
    """A Robin side and an enforced side on one mesh, batched with a plain problem and a
    second enforced one: each sub-problem equals its own solve bit for bit and the oracle."""
    from pyoomph_amd import workloads as wl
    from pyoomph_amd.problem import Problem
    mp1, op1 = build(7, "C2", "left", 1.0, enf=(1.0, -0.2), left=0.1,
                     flux=[(("top",), 2.0, 0.5)])
    mp2, op2 = build(5, "C1", "left", 0.7, enf=(2.0, 0.3), left=0.0)
    plain = wl.bratu(9, "C2", 1.1)
    parts = [mp1, plain, mp2]
    B = solver(Problem(parts))
    B.solve()
    xb = B.get_current_dofs()
    for i, mp in enumerate(parts):
        A = solver(Problem(mp))
        A.solve()
        assert np.array_equal(B.sub_dofs(i, xb), A.get_current_dofs()), i
    for i, op in ((0, op1), (2, op2)):
        assert rel(B.sub_dofs(i, xb), on.newton_solve(op, np.zeros(op.ndof), tol=1e-12)) <= 1e-8


def test_saddle_point_needs_ilu0(gpu):
    from pyoomph_amd.problem import Problem
    mp, _ = build(4, "C1", "left", 0.0)
    P = Problem(mp)
    P.preconditioner = "amg"
    with pytest.raises(ValueError, match="saddle point"):
        P.solve()


def _constraint_pair(space, split, form, coef, cexpr, lam=1.1, N=6):
    """The product problem with a generated constraint form on the right side and the
    oracle with the same constraint expression."""
    from pyoomph_amd.equations import BratuEquation
    from pyoomph_amd.meshes import RectangularQuadMesh
    from pyoomph_amd.problem import MeshProblem
    mp = MeshProblem(RectangularQuadMesh(N, split_in_tris=split, space=space),
                     BratuEquation("u", space=space, lambda_=lam), dirichlet=[(("left",), {0: 1.0})],
                     enforced=[(("right",), coef)], enforced_form=form)
    _, ref = build(N, space, split, lam)
    faces = of.boundary_faces(ref.mesh, ["right"])
    op = oa.OracleProblem(ref.mesh, oa.FORM_BRATU, 2, ref.pinned, ref.bc, [lam], enforced=faces, enf_expr=cexpr,
                          enf_coef=np.tile(np.asarray(coef, float), (len(faces), 1)))
    assert np.array_equal(mp.eqn, op.eqn)
    return mp, op


@pytest.mark.parametrize("space,split", CASES)
def test_generated_linear_constraint_equals_the_built_in_faces_bitwise(gpu, space, split):
    from pyoomph_amd.codegen.face import ConstraintForm
    from pyoomph_amd.problem import Problem
    form = ConstraintForm("gen_lin_constraint", lambda s: s.c[0] * s.u + s.c[1], ncoef=2)
    A = Problem(build(6, space, split, 1.5, enf=(1.7, -0.4))[0])
    mp, _ = _constraint_pair(space, split, form, (1.7, -0.4), None, lam=1.5)
    B = Problem(mp)
    x = np.random.default_rng(1).uniform(-0.3, 0.3, A.ndof)
    A.set_current_dofs(x)
    B.set_current_dofs(x)
    RA, JA = A.assemble_jacobian()
    RB, JB = B.assemble_jacobian()
    assert np.array_equal(RA, RB)
    assert np.array_equal(JA.indptr, JB.indptr) and np.array_equal(JA.indices, JB.indices)
    assert np.array_equal(JA.data, JB.data)
    assert np.array_equal(A.get_residuals(dparameter="lambda"), B.get_residuals(dparameter="lambda"))


@pytest.mark.parametrize("space,split", CASES)
def test_generated_nonlinear_constraint_matches_the_oracle(gpu, space, split):
    """EnforcedBC(u = u^3 + c0 y - lambda c1 u): assembly and dR/dlambda within 1e-12 of
    the oracle's constraint expression, Newton (GMRES + ILU(0), multipliers last) within
    1e-8 of the oracle's."""
    from pyoomph_amd.codegen.face import ConstraintForm
    from pyoomph_amd.problem import Problem
    lam, coef = 1.1, (0.5, 2.0)
    form = ConstraintForm("gen_cubic_constraint", lambda s: s.u ** 3 + s.c[0] * s.x[1] - s.p[0] * s.c[1] * s.u,
                          ncoef=2, nparams=1, global_param=0)
    cexpr = (lambda u, x0, x1, C: u ** 3 + C[:, 0] * x1 - lam * C[:, 1] * u,
             lambda u, x0, x1, C: 3.0 * u ** 2 - lam * C[:, 1],
             lambda u, x0, x1, C: -(C[:, 1] * u),
             lambda u, x0, x1, C: -C[:, 1] + 0.0 * u)
    mp, op = _constraint_pair(space, split, form, coef, cexpr, lam=lam)
    P = solver(Problem(mp))
    x = np.random.default_rng(4).uniform(-0.3, 0.3, P.ndof)
    P.set_current_dofs(x)
    R, J = P.assemble_jacobian()
    r, rs, col, val = op.get_jacobian(x)
    assert np.array_equal(J.indptr, rs) and np.array_equal(J.indices, col)
    assert rel(J.data, val) <= 1e-12 and rel(R, r) <= 1e-12
    assert rel(P.get_residuals(dparameter="lambda"), op.get_residuals(x, deriv=1)) <= 1e-12
    P.dofs.zero_()
    st = P.solve()
    assert st[0]["converged"]
    xo = on.newton_solve(op, np.zeros(op.ndof), tol=1e-11)
    assert rel(P.get_current_dofs(), xo) <= 1e-8


@pytest.mark.parametrize("space,split", CASES)
def test_hessian_products_with_generated_faces(gpu, space, split):
    """Bratu with a generated cubic flux on the top side and a generated cubic constraint on
    the right: the plain and the transposed Hessian-vector products (bulk and both face
    kinds) against central differences of the Jacobian, H v = d/de J(x + e v) y and
    H^T-product v = d/de J(x + e v)^T y."""
    from pyoomph_amd.codegen.face import ConstraintForm, FaceFluxForm
    from pyoomph_amd.equations import BratuEquation
    from pyoomph_amd.meshes import RectangularQuadMesh
    from pyoomph_amd.problem import MeshProblem, Problem
    flux = FaceFluxForm("gen_hv_flux", lambda s: s.c[0] * s.u ** 3 + s.c[1] * s.n[1], ncoef=2)
    con = ConstraintForm("gen_hv_constraint", lambda s: s.u ** 3 + s.c[0] * s.x[1] - s.p[0] * s.u, ncoef=1,
                         nparams=1, global_param=0)
    mp = MeshProblem(RectangularQuadMesh(5, split_in_tris=split, space=space),
                     BratuEquation("u", space=space, lambda_=0.9), dirichlet=[(("left",), {0: 0.5})],
                     flux=[(("top",), (0.8, -0.2))], flux_form=flux,
                     enforced=[(("right",), (0.3,))], enforced_form=con)
    P = Problem(mp)
    rng = np.random.default_rng(7)
    x, y, v = (rng.uniform(-0.4, 0.4, P.ndof) for _ in range(3))
    P.set_current_dofs(x)
    H = P.assemble_hessian_vector_product(y)
    Ht = P.assemble_hessian_vector_product(y, transposed=True)
    eps = 1e-5
    P.set_current_dofs(x + eps * v)
    Jp = P.assemble_jacobian()[1]
    P.set_current_dofs(x - eps * v)
    Jm = P.assemble_jacobian()[1]
    fd = (Jp @ y - Jm @ y) / (2 * eps)
    fdt = (Jp.T @ y - Jm.T @ y) / (2 * eps)
    assert np.max(np.abs(H @ v - fd)) <= 1e-7 * np.max(np.abs(fd))
    assert np.max(np.abs(Ht @ v - fdt)) <= 1e-7 * np.max(np.abs(fdt))
