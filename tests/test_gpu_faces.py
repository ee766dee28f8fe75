"""GPU parity of the boundary-face flux elements (NeumannBC(u=alpha*u + beta) on the
faces of P1 / P2 triangles) against the oracle: CSR pattern bit-exact, values and
residuals within 1e-12, Newton solutions within 1e-8, the exact linear solution of a
Laplace problem with a flux side, and batch = alone bitwise."""
import numpy as np
import pytest

from oracle import assembly as oa
from oracle import faces as of
from oracle import meshes as om
from oracle import newton as on
from oracle import timestepping as ots

pytestmark = pytest.mark.gpu

FLUX, FIXED = ("right", "top"), ("left", "bottom")


def build(N, space, lam, alpha, beta, split="left"):
    from pyoomph_amd.equations import BratuEquation
    from pyoomph_amd.meshes import RectangularQuadMesh
    from pyoomph_amd.problem import MeshProblem, Problem
    mesh = RectangularQuadMesh(N, split_in_tris=split, space=space)
    mp = MeshProblem(mesh, BratuEquation("u", space=space, lambda_=lam), dirichlet=[(FIXED, {0: 0.0})],
                     flux=[(FLUX, alpha, beta)])
    m = om.rectangular_quad_mesh(N, split, space)
    op = oa.OracleProblem(m, oa.FORM_BRATU, 1, m.on_boundary(set(FIXED))[:, None], np.zeros((m.nnode, 1)),
                          [lam], faces=of.boundary_faces(m, list(FLUX)), face_alpha=alpha, face_beta=beta)
    assert np.array_equal(mp.eqn, op.eqn) and np.array_equal(mp.faces, op.faces)
    return Problem(mp), op


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("space,split", [("C1", "left"), ("C2", "left"), ("C2", False)])
@pytest.mark.parametrize("alpha,beta", [(0.0, 0.6), (1.3, -0.4)])
def test_flux_assembly_parity(gpu, space, split, alpha, beta):
    P, op = build(7, space, 1.2, alpha, beta, split)
    for seed in (1, 2):
        x = np.random.default_rng(seed).uniform(-0.4, 0.4, P.ndof)
        P.set_current_dofs(x)
        R, J = P.assemble_jacobian()
        r, rs, col, val = op.get_jacobian(x)
        assert np.array_equal(J.indptr, rs) and np.array_equal(J.indices, col)
        assert rel(J.data, val) <= 1e-12 and rel(R, r) <= 1e-12
        assert np.array_equal(P.get_residuals(), R)
        d = P.get_residuals(dparameter="lambda")
        assert rel(d, op.get_residuals(x, deriv=1)) <= 1e-12


def test_flux_mass_matrix_has_no_face_term(gpu):
    P, op = build(6, "C2", 1.0, 2.0, 0.5)
    x = np.random.default_rng(3).uniform(-0.2, 0.2, P.ndof)
    P.set_current_dofs(x)
    M = P.assemble_mass_matrix()
    op.set_time(ots.bdf1_weights(1.0), x, x)
    _, rs, col, val = op.get_mass_matrix(x)
    op.set_time(None, None, None)
    assert np.array_equal(M.indptr, rs) and np.array_equal(M.indices, col)
    assert np.array_equal(M.data, val)


@pytest.mark.parametrize("space,split", [("C1", "left"), ("C2", "left"), ("C2", False)])
@pytest.mark.parametrize("alpha,beta", [(0.0, 0.7), (2.0, 0.7)])
def test_laplace_with_flux_side_has_linear_solution(gpu, space, split, alpha, beta):
    """u = 0 on the left, flux on the right only: u = -beta / (1 + alpha) * x."""
    from pyoomph_amd.equations import BratuEquation
    from pyoomph_amd.meshes import RectangularQuadMesh
    from pyoomph_amd.problem import MeshProblem, Problem
    mp = MeshProblem(RectangularQuadMesh(9, split_in_tris=split, space=space),
                     BratuEquation("u", space=space, lambda_=0.0), dirichlet=[(("left",), {0: 0.0})],
                     flux=[(("right",), alpha, beta)])
    P = Problem(mp)
    P.krylov, P.preconditioner = "cg", "jacobi"
    P.newton_solver_tolerance = 1e-12
    st = P.solve()
    assert st[0]["converged"]
    u = mp.nodal_values(P.get_current_dofs())[:, 0]
    c = -beta / (1.0 + alpha)
    # Q2: the reference's Gauss<2,3> positive knot is mistyped (~6e-9), see the oracle test
    assert np.max(np.abs(u - c * mp.mesh.coords[:, 0])) <= (1e-10 if split else 5e-9)


@pytest.mark.parametrize("opts", [dict(krylov="cg", preconditioner="jacobi"),
                                  dict(krylov="cg", preconditioner="amg"),
                                  dict(krylov="gmres", preconditioner="ilu0")])
@pytest.mark.parametrize("space,split", [("C1", "left"), ("C2", "left"), ("C2", False)])
def test_flux_newton_matches_oracle(gpu, opts, space, split):
    P, op = build(14, space, 1.8, 0.8, 0.3, split)
    for k, v in opts.items():
        setattr(P, k, v)
    st = P.solve()
    assert st[0]["converged"] and st[0]["linear_failures"] == 0
    xo = on.newton_solve(op, np.zeros(op.ndof), tol=1e-8)
    assert rel(P.get_current_dofs(), xo) <= 1e-8


def test_flux_batch_equals_alone_bitwise(gpu):
    from pyoomph_amd import workloads as wl
    from pyoomph_amd.problem import Problem
    parts = [lambda: wl.bratu_flux(11, "C2", 1.1, 0.5, 0.2), lambda: wl.bratu(9, "C1", 2.0),
             lambda: wl.bratu_flux(13, "C1", 0.6, 0.0, -0.3), lambda: wl.bratu_flux(6, "C2", 2.2, 3.0, 0.0)]
    batch = Problem([f() for f in parts])
    batch.krylov, batch.preconditioner = "cg", "amg"
    batch.solve()
    xb = batch.get_current_dofs()
    for i, f in enumerate(parts):
        alone = Problem(f())
        alone.krylov, alone.preconditioner = "cg", "amg"
        alone.solve()
        assert np.array_equal(alone.get_current_dofs(), batch.sub_dofs(i, xb)), i


@pytest.mark.parametrize("space", ["C1", "C2"])
def test_several_fluxes_on_one_mesh(gpu, space):
    """Neumann on the right, Robin on the top, u = 0 on the left and bottom: per-face flux
    coefficients against the oracle (assembly and Newton)."""
    from pyoomph_amd.equations import BratuEquation
    from pyoomph_amd.meshes import RectangularQuadMesh
    from pyoomph_amd.problem import MeshProblem, Problem
    N, lam = 10, 1.4
    mesh = RectangularQuadMesh(N, split_in_tris="left", space=space)
    mp = MeshProblem(mesh, BratuEquation("u", space=space, lambda_=lam), dirichlet=[(FIXED, {0: 0.0})],
                     flux=[(("right",), 0.0, 0.45), (("top",), 1.7, -0.25)])
    m = om.rectangular_quad_mesh(N, "left", space)
    Fr, Ft = of.boundary_faces(m, ["right"]), of.boundary_faces(m, ["top"])
    op = oa.OracleProblem(m, oa.FORM_BRATU, 1, m.on_boundary(set(FIXED))[:, None], np.zeros((m.nnode, 1)), [lam],
                          faces=np.concatenate([Fr, Ft]),
                          face_alpha=np.r_[np.zeros(len(Fr)), np.full(len(Ft), 1.7)],
                          face_beta=np.r_[np.full(len(Fr), 0.45), np.full(len(Ft), -0.25)])
    assert np.array_equal(mp.faces, op.faces)
    P = Problem(mp)
    x = np.random.default_rng(6).uniform(-0.3, 0.3, P.ndof)
    P.set_current_dofs(x)
    R, J = P.assemble_jacobian()
    r, rs, col, val = op.get_jacobian(x)
    assert np.array_equal(J.indptr, rs) and np.array_equal(J.indices, col)
    assert rel(J.data, val) <= 1e-12 and rel(R, r) <= 1e-12
    P.set_current_dofs(np.zeros(P.ndof))
    P.krylov, P.preconditioner = "cg", "amg"
    st = P.solve()
    assert st[0]["converged"]
    assert rel(P.get_current_dofs(), on.newton_solve(op, np.zeros(op.ndof), tol=1e-8)) <= 1e-8


@pytest.mark.parametrize("alpha,beta", [(0.0, 0.6), (1.3, -0.4)])
def test_brick_flux_parity_and_newton(gpu, alpha, beta):
    """Faces of Q2 bricks (9-node quads in 3D): assembly against the oracle, then Newton."""
    from pyoomph_amd.equations import BratuEquation
    from pyoomph_amd.meshes import CuboidBrickMesh
    from pyoomph_amd.problem import MeshProblem, Problem
    N, lam = 4, 1.1
    mp = MeshProblem(CuboidBrickMesh(N), BratuEquation("u", space="C2", lambda_=lam),
                     dirichlet=[(("left", "bottom"), {0: 0.0})],
                     flux=[(("right", "top"), alpha, beta), (("front",), 0.5 * alpha, 0.3)])
    m = om.cuboid_brick_mesh(N)
    Fa, Fb = of.boundary_faces(m, ["right", "top"]), of.boundary_faces(m, ["front"])
    op = oa.OracleProblem(m, oa.FORM_BRATU, 1, m.on_boundary({"left", "bottom"})[:, None],
                          np.zeros((m.nnode, 1)), [lam], faces=np.concatenate([Fa, Fb]),
                          face_alpha=np.r_[np.full(len(Fa), alpha), np.full(len(Fb), 0.5 * alpha)],
                          face_beta=np.r_[np.full(len(Fa), beta), np.full(len(Fb), 0.3)])
    assert np.array_equal(mp.eqn, op.eqn) and np.array_equal(mp.faces, op.faces)
    P = Problem(mp)
    x = np.random.default_rng(9).uniform(-0.3, 0.3, P.ndof)
    P.set_current_dofs(x)
    R, J = P.assemble_jacobian()
    r, rs, col, val = op.get_jacobian(x)
    assert np.array_equal(J.indptr, rs) and np.array_equal(J.indices, col)
    assert rel(J.data, val) <= 1e-12 and rel(R, r) <= 1e-12
    P.set_current_dofs(np.zeros(P.ndof))
    P.krylov, P.preconditioner = "cg", "amg"
    st = P.solve()
    assert st[0]["converged"]
    assert rel(P.get_current_dofs(), on.newton_solve(op, np.zeros(op.ndof), tol=1e-8)) <= 1e-8


def test_brick_laplace_with_flux_side_has_linear_solution(gpu):
    from pyoomph_amd.equations import BratuEquation
    from pyoomph_amd.meshes import CuboidBrickMesh
    from pyoomph_amd.problem import MeshProblem, Problem
    mp = MeshProblem(CuboidBrickMesh(3), BratuEquation("u", space="C2", lambda_=0.0),
                     dirichlet=[(("left",), {0: 0.0})], flux=[(("right",), 2.0, 0.7)])
    P = Problem(mp)
    P.krylov, P.preconditioner = "cg", "jacobi"
    P.newton_solver_tolerance = 1e-12
    assert P.solve()[0]["converged"]
    u = mp.nodal_values(P.get_current_dofs())[:, 0]
    # the reference's mistyped Gauss<2,3> knot on the faces bounds the exactness (~1e-8)
    assert np.max(np.abs(u - (-0.7 / 3.0) * mp.mesh.coords[:, 0])) <= 1e-7


@pytest.mark.parametrize("alpha,beta", [(0.0, (0.1, -0.05)), (0.8, (0.0, 0.2))])
def test_taylor_hood_traction_parity_and_newton(gpu, alpha, beta):
    """Taylor-Hood with a traction face (NeumannBC on the velocity components) on the right
    of the driven cavity: assembly against the oracle, then Newton (GMRES + ILU(0))."""
    from pyoomph_amd.equations import NavierStokesEquations
    from pyoomph_amd.meshes import RectangularQuadMesh
    from pyoomph_amd.problem import MeshProblem, Problem
    from tests.test_oracle_faces import _ns_traction_problem
    N = 6
    mp = MeshProblem(RectangularQuadMesh(N, split_in_tris=False, space="C2"),
                     NavierStokesEquations(dynamic_viscosity=0.05, mass_density=1.0),
                     dirichlet=[(("top",), {0: 1.0, 1: 0.0}), (("left", "bottom"), {0: 0.0, 1: 0.0})],
                     pin_extra=[(0, 2, 0.0)], flux=[(("right",), alpha, beta)])
    _, op = _ns_traction_problem(N, alpha, beta)
    assert np.array_equal(mp.eqn, op.eqn) and np.array_equal(mp.faces, op.faces)
    P = Problem(mp)
    x = np.random.default_rng(4).uniform(-0.3, 0.3, P.ndof)
    P.set_current_dofs(x)
    R, J = P.assemble_jacobian()
    r, rs, col, val = op.get_jacobian(x)
    assert np.array_equal(J.indptr, rs) and np.array_equal(J.indices, col)
    assert np.array_equal(J.data, val) and np.array_equal(R, r)
    P.set_current_dofs(np.zeros(P.ndof))
    P.krylov, P.preconditioner, P.gmres_restart = "gmres", "ilu0", 100
    P.newton_solver_tolerance = 1e-10
    st = P.solve()
    assert st[0]["converged"]
    assert rel(P.get_current_dofs(), on.newton_solve(op, np.zeros(op.ndof), tol=1e-10)) <= 1e-8
