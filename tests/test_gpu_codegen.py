"""Run-time generated element kernels (pyoomph_amd/codegen) on the GPU.

* The generated Poisson, Bratu and reaction-diffusion kernels give the compiled kernels'
  FP64 residuals, Jacobian values and dR/dlambda bit for bit (same operations in the same
  order), steady and with the time derivative; FP32 within 1e-6. Newton solves through the
  generated kernels give the compiled path's bits and the oracle's solution (1e-8).
* A form with no compiled counterpart, -div((1 + k u^2) grad u) = s0 with u = 0 on the unit
  square: by the Kirchhoff transform v = u + k u^3 / 3 solves -lap v = s0, whose solution is
  the double sine series 16 s0 / (pi^4 m n (m^2 + n^2)) sin(m pi x) sin(n pi y), m, n odd;
  the P2 solution converges to it, its Jacobian equals central differences of its residual,
  and Newton converges quadratically.
"""
import numpy as np
import pytest

from oracle import newton as on
from pyoomph_amd import workloads as wl
from pyoomph_amd.codegen import library
from pyoomph_amd.codegen.runtime import GeneratedEquation
from pyoomph_amd.meshes import CuboidBrickMesh, RectangularQuadMesh
from tests.oracle_bridge import oracle_problem

pytestmark = pytest.mark.gpu

ALL2D = ("left", "right", "bottom", "top")
ALL3D = ALL2D + ("back", "front")


def pair(kind, **kw):
    """(compiled-form MeshProblem, generated-form MeshProblem) on the same mesh."""
    from pyoomph_amd.problem import MeshProblem
    if kind == "bratu":
        a = wl.bratu(kw["N"], kw["space"], kw["lam"])
        mesh = RectangularQuadMesh(kw["N"], split_in_tris="left", space=kw["space"])
        eq = GeneratedEquation(library.bratu(), space=kw["space"], params=[kw["lam"]])
        return a, MeshProblem(mesh, eq, dirichlet=[(ALL2D, {0: 0.0})])
    if kind == "poisson":
        a = wl.poisson_gaussians(kw["N"], kw.get("seed", 0))
        mesh = RectangularQuadMesh(kw["N"], split_in_tris="left", space="C2")
        eq = GeneratedEquation(library.poisson(8), space="C2", params=a.equations.params())
        return a, MeshProblem(mesh, eq, dirichlet=[(ALL2D, {0: 0.0})])
    if kind == "rd":
        a = wl.reaction_diffusion_cube(kw["N"], kw.get("seed", 0))
        mesh = CuboidBrickMesh(kw["N"])
        eq = GeneratedEquation(library.reaction_diffusion(), space="C2", aux=a.equations.source_field)
        return a, MeshProblem(mesh, eq, dirichlet=[(ALL3D, {0: 0.0})])
    raise KeyError(kind)


CASES = [("bratu", dict(N=9, space="C1", lam=1.5)), ("bratu", dict(N=7, space="C2", lam=2.5)),
         ("poisson", dict(N=8, seed=3)), ("rd", dict(N=3))]


@pytest.mark.parametrize("kind,kw", CASES)
def test_generated_kernels_equal_compiled_bitwise(gpu, kind, kw):
    from pyoomph_amd.problem import Problem
    mp_a, mp_b = pair(kind, **kw)
    A, B = Problem(mp_a), Problem(mp_b)
    assert A.ndof == B.ndof and A.nnz == B.nnz
    rng = np.random.default_rng(1)
    for _ in range(2):
        x = rng.uniform(-0.4, 0.4, A.ndof)
        A.set_current_dofs(x)
        B.set_current_dofs(x)
        ra, va = A.assemble_device("f64")
        rb, vb = B.assemble_device("f64")
        assert np.array_equal(ra.cpu().numpy(), rb.cpu().numpy())
        assert np.array_equal(va.cpu().numpy(), vb.cpu().numpy())
        if kind == "bratu":
            da, dva = A.assemble_device("f64", dparameter="lambda")
            db, dvb = B.assemble_device("f64", dparameter="lambda")
            assert np.array_equal(da.cpu().numpy(), db.cpu().numpy())
            assert np.array_equal(dva.cpu().numpy(), dvb.cpu().numpy())
        # FP32 kernels (the compiled ones contract to FMA, the generated ones do not)
        r32a, v32a = A.assemble_device("f32")
        r32b, v32b = B.assemble_device("f32")
        scale = np.max(np.abs(va.cpu().numpy()))
        assert np.max(np.abs(v32a.cpu().numpy().astype(float) - v32b.cpu().numpy())) <= 1e-6 * scale
        # the time derivative (transient instantiations)
        h1, h2 = (A.torch.tensor(rng.uniform(-0.3, 0.3, A.ndof), device=A.device) for _ in range(2))
        w = A.bdf_weights(0.1, 0.07)
        A.set_time_state(w, h1, h2)
        B.set_time_state(w, h1, h2)
        ta, tva = A.assemble_device("f64")
        tb, tvb = B.assemble_device("f64")
        A.set_time_state(None, None, None)
        B.set_time_state(None, None, None)
        assert np.array_equal(ta.cpu().numpy(), tb.cpu().numpy())
        assert np.array_equal(tva.cpu().numpy(), tvb.cpu().numpy())


@pytest.mark.parametrize("kind,kw,precond", [("bratu", dict(N=20, space="C2", lam=2.0), "amg"),
                                             ("bratu", dict(N=17, space="C1", lam=0.9), "jacobi"),
                                             ("poisson", dict(N=16, seed=2), "amg"),
                                             ("rd", dict(N=4), "ilu0")])
def test_generated_newton_equals_compiled_and_oracle(gpu, kind, kw, precond):
    from pyoomph_amd.problem import Problem
    mp_a, mp_b = pair(kind, **kw)
    out = []
    for mp in (mp_a, mp_b):
        P = Problem(mp)
        P.krylov, P.preconditioner = "cg", precond
        st = P.solve()
        assert st[0]["converged"]
        out.append(P.get_current_dofs())
    assert np.array_equal(out[0], out[1])
    okw = dict(kw)
    if kind == "rd":
        okw["f"] = mp_a.equations.source_field
    op = oracle_problem(kind, **okw)
    xo = on.newton_solve(op, np.zeros(op.ndof))
    assert np.max(np.abs(out[1] - xo)) <= 1e-8 * np.max(np.abs(xo))


def _series_v(x, y, s0, M=301):
    """Solution of -lap v = s0 on the unit square, v = 0 on the boundary."""
    m = np.arange(1, M + 1, 2, dtype=np.float64)
    v = np.zeros_like(x)
    sx = np.sin(np.pi * np.outer(x, m))          # [npts, nm]
    sy = np.sin(np.pi * np.outer(y, m))
    coef = 16.0 * s0 / (np.pi ** 4 * np.outer(m, m) * (m[:, None] ** 2 + m[None, :] ** 2))
    v = np.einsum("pi,ij,pj->p", sx, coef, sy)
    return v


def _solve_nd(N, k, s0):
    from pyoomph_amd.problem import MeshProblem, Problem
    mesh = RectangularQuadMesh(N, split_in_tris="left", space="C2")
    eq = GeneratedEquation(library.nonlinear_diffusion(), space="C2", params=[k, s0])
    mp = MeshProblem(mesh, eq, dirichlet=[(ALL2D, {0: 0.0})])
    P = Problem(mp)
    P.krylov, P.preconditioner, P.gmres_restart = "gmres", "ilu0", 100
    st = P.solve()
    return mp, P, st


def test_nonlinear_diffusion_matches_the_kirchhoff_transform(gpu):
    k, s0 = 2.0, 8.0
    errs = []
    for N in (8, 16):
        mp, P, st = _solve_nd(N, k, s0)
        assert st[0]["converged"] and st[0]["newton_iters"] <= 6
        # quadratic convergence: each residual at most ~ C * previous^2 once small
        h = st[0]["history"]
        assert h[-1] < 1e-8 and any(h[i + 1] < 10 * h[i] ** 2 for i in range(len(h) - 1) if h[i] < 1e-2)
        U = mp.nodal_values(P.get_current_dofs())[:, 0]
        c = mp.mesh.coords
        v = _series_v(c[:, 0], c[:, 1], s0)
        u = v.copy()
        for _ in range(50):     # u + k u^3 / 3 = v, pointwise Newton
            u = u - (u + k * u ** 3 / 3 - v) / (1 + k * u * u)
        errs.append(np.max(np.abs(U - u)) / np.max(np.abs(u)))
    assert errs[1] < 2e-3, errs
    assert errs[1] < errs[0] / 3.0, errs     # converging with the mesh


def test_nonlinear_diffusion_jacobian_is_the_residual_derivative(gpu):
    mp, P, st = _solve_nd(6, 2.0, 8.0)
    rng = np.random.default_rng(4)
    x = rng.uniform(-0.5, 0.5, P.ndof)
    P.set_current_dofs(x)
    R, J = P.assemble_jacobian()
    d = rng.normal(size=P.ndof)
    eps = 1e-6
    P.set_current_dofs(x + eps * d)
    rp = P.get_residuals()
    P.set_current_dofs(x - eps * d)
    rm = P.get_residuals()
    fd = (rp - rm) / (2 * eps)
    assert np.max(np.abs(J @ d - fd)) <= 1e-6 * np.max(np.abs(fd))
    assert not np.allclose(J.toarray(), J.toarray().T)   # dF/du makes it non-symmetric


def _th_pair(N, Re):
    """(hand-written Navier-Stokes MeshProblem, generated Navier-Stokes MeshProblem)."""
    from pyoomph_amd.problem import MeshProblem
    a = wl.lid_driven_cavity(N, Re=Re)
    mesh = RectangularQuadMesh(N, split_in_tris=False, space="C2")
    eq = GeneratedEquation(library.navier_stokes(), space="C2", params=[1.0 / Re, 1.0])
    bcs = [(("top",), {0: 1.0, 1: 0.0}), (("left", "right", "bottom"), {0: 0.0, 1: 0.0})]
    return a, MeshProblem(mesh, eq, dirichlet=bcs, pin_extra=[(0, 2, 0.0)])


def test_generated_taylor_hood_kernel_equals_the_hand_written_one(gpu):
    """Generated Navier-Stokes (Taylor-Hood layout) against the hand-written kernel: the
    residual bit for bit (same operations in the same order), the Jacobian to 1e-13 (the
    generator groups the viscous and convection terms per derivative), steady and with the
    BDF2 time term; the mass matrix; FP32 within 1e-5."""
    from pyoomph_amd.problem import Problem
    mp_a, mp_b = _th_pair(6, 80.0)
    A, B = Problem(mp_a), Problem(mp_b)
    assert A.ndof == B.ndof and A.nnz == B.nnz
    assert np.array_equal(A.structural_pattern()[1], B.structural_pattern()[1])
    rng = np.random.default_rng(3)
    for transient in (False, True):
        x = rng.uniform(-0.4, 0.4, A.ndof)
        A.set_current_dofs(x)
        B.set_current_dofs(x)
        if transient:
            h1, h2 = (A.torch.tensor(rng.uniform(-0.3, 0.3, A.ndof), device=A.device) for _ in range(2))
            w = A.bdf_weights(0.05, 0.04)
            A.set_time_state(w, h1, h2)
            B.set_time_state(w, h1, h2)
        ra, va = A.assemble_device("f64")
        rb, vb = B.assemble_device("f64")
        ra, rb, va, vb = (t.cpu().numpy() for t in (ra, rb, va, vb))
        assert np.array_equal(ra, rb), transient
        assert np.max(np.abs(va - vb)) <= 1e-13 * np.max(np.abs(va))
        if transient:
            Ma, Mb = A.assemble_mass_matrix(), B.assemble_mass_matrix()
            assert np.max(np.abs((Ma - Mb).toarray())) <= 1e-14 * np.max(np.abs(Ma.toarray()))
            A.set_time_state(None, None, None)
            B.set_time_state(None, None, None)
    _, v64 = B.assemble_device("f64")          # steady, at the last dofs
    r32, v32 = B.assemble_device("f32")
    v64 = v64.cpu().numpy()
    assert np.max(np.abs(v32.cpu().numpy().astype(float) - v64)) <= 1e-5 * np.max(np.abs(v64))


def test_generated_taylor_hood_newton_reaches_the_oracle(gpu):
    from pyoomph_amd.problem import Problem
    _, mp = _th_pair(16, 100.0)
    P = Problem(mp)
    P.krylov, P.preconditioner, P.gmres_restart = "gmres", "block", 100
    P.newton_solver_tolerance, P.gmres_max_iters = 1e-12, 5000
    st = P.solve()
    assert st[0]["converged"] and st[0]["linear_failures"] == 0
    op = oracle_problem("ns", N=16, Re=100.0)
    xo = on.converged_solve(op, np.zeros(op.ndof), tol=1e-12)
    assert np.max(np.abs(P.get_current_dofs() - xo)) <= 1e-8 * np.max(np.abs(xo))


def test_generated_stokes_brinkman(gpu):
    """A Taylor-Hood form with no hand-written counterpart: Stokes flow with a Brinkman drag
    k u and a body force; the problem is linear, so Newton converges at once, the
    Jacobian equals central differences of the residual, and with no body force and no
    wall motion the solution is zero."""
    from pyoomph_amd.problem import MeshProblem, Problem
    mesh = RectangularQuadMesh(8, split_in_tris=False, space="C2")
    eq = GeneratedEquation(library.stokes_brinkman(), space="C2", params=[0.3, 2.0, 1.5])
    walls = [(("left", "right", "bottom", "top"), {0: 0.0, 1: 0.0})]
    mp = MeshProblem(mesh, eq, dirichlet=walls, pin_extra=[(0, 2, 0.0)])
    P = Problem(mp)
    P.krylov, P.preconditioner, P.gmres_restart = "gmres", "ilu0", 200
    P.newton_solver_tolerance = 1e-11
    st = P.solve()
    assert st[0]["converged"] and st[0]["newton_iters"] <= 2
    rng = np.random.default_rng(6)
    x = rng.uniform(-0.5, 0.5, P.ndof)
    P.set_current_dofs(x)
    R, J = P.assemble_jacobian()
    d = rng.normal(size=P.ndof)
    eps = 1e-6
    P.set_current_dofs(x + eps * d)
    rp = P.get_residuals()
    P.set_current_dofs(x - eps * d)
    rm = P.get_residuals()
    assert np.max(np.abs(J @ d - (rp - rm) / (2 * eps))) <= 1e-7 * np.max(np.abs(J @ d))
    # body force g acts on u_y only: the pressure -g y balances it and u = 0 is the discrete
    # solution (up to the linear solver's tolerance)
    P.dofs.zero_()
    P.solve()
    U = mp.nodal_values(P.get_current_dofs())
    assert np.max(np.abs(U[:, :2])) <= 1e-8
    corner = np.isin(np.arange(mesh.nnode), mesh.elements[:, [0, 2, 6, 8]].reshape(-1))
    y = mesh.coords[corner, 1]
    assert np.max(np.abs(U[corner, 2] - (-1.5 * y))) <= 1e-7


# ---- generated boundary-face forms --------------------------------------------------------
def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _face_pair(space, split, flux_form=None, flux=None, lam=1.0, N=7):
    from pyoomph_amd.equations import BratuEquation
    from pyoomph_amd.problem import MeshProblem
    mesh = RectangularQuadMesh(N, split_in_tris=split, space=space)
    return MeshProblem(mesh, BratuEquation("u", space=space, lambda_=lam), dirichlet=[(("left",), {0: 0.0})],
                       flux=flux, flux_form=flux_form)


FACE_CASES = [("C1", "left"), ("C2", "left"), ("C2", False)]


@pytest.mark.parametrize("space,split", FACE_CASES)
def test_generated_linear_flux_equals_the_built_in_faces_bitwise(gpu, space, split):
    """NeumannBC(u = c0*u + c1) generated from its expression reproduces the library's
    flux faces bit for bit (residual, Jacobian pattern and values, dR/dlambda)."""
    from pyoomph_amd.codegen.face import FaceFluxForm
    from pyoomph_amd.problem import Problem
    lin = FaceFluxForm("gen_robin", lambda s: s.c[0] * s.u + s.c[1], ncoef=2)
    A = Problem(_face_pair(space, split, flux=[(("right", "top"), 1.3, -0.4)]))
    B = Problem(_face_pair(space, split, flux_form=lin, flux=[(("right", "top"), (1.3, -0.4))]))
    for seed in (1, 2):
        x = np.random.default_rng(seed).uniform(-0.3, 0.3, A.ndof)
        A.set_current_dofs(x)
        B.set_current_dofs(x)
        RA, JA = A.assemble_jacobian()
        RB, JB = B.assemble_jacobian()
        assert np.array_equal(RA, RB)
        assert np.array_equal(JA.indptr, JB.indptr) and np.array_equal(JA.indices, JB.indices)
        assert np.array_equal(JA.data, JB.data)
        assert np.array_equal(A.get_residuals(dparameter="lambda"), B.get_residuals(dparameter="lambda"))


def _cubic_form():
    from pyoomph_amd.codegen import expr as E
    from pyoomph_amd.codegen.face import FaceFluxForm
    return FaceFluxForm("gen_cubic_flux", lambda s: (s.c[0] * s.u ** 3 + s.c[1] * s.x[1]) - s.p[0] * E.exp(s.u),
                        ncoef=2, nparams=1, global_param=0)


def _cubic_oracle(space, split, lam, c, N=7):
    from oracle import assembly as oa
    from oracle import faces as of
    from oracle import meshes as om
    m = om.rectangular_quad_mesh(N, split, space)
    faces = of.boundary_faces(m, ["right", "top"])
    C = np.tile(np.asarray(c, dtype=np.float64), (len(faces), 1))
    flux = lambda u, x0, x1, C: (C[:, 0] * ((u * u) * u) + C[:, 1] * x1) - lam * np.exp(u)  # noqa: E731
    dflux = lambda u, x0, x1, C: C[:, 0] * (3.0 * (u * u)) - lam * np.exp(u)  # noqa: E731
    dp = lambda u, x0, x1, C: -np.exp(u)  # noqa: E731
    pinned = m.on_boundary(["left"])[:, None]
    return oa.OracleProblem(m, oa.FORM_BRATU, 1, pinned, np.zeros((m.nnode, 1)), [lam], faces=faces,
                            face_expr=(flux, dflux, dp, dp), face_coef=C)


@pytest.mark.parametrize("space,split", FACE_CASES)
def test_generated_nonlinear_flux_matches_the_oracle(gpu, space, split):
    """NeumannBC(u = c0 u^3 + c1 y - lambda e^u): assembly, dR/dlambda and dJ/dlambda within
    1e-12 of the oracle's expression faces (exponentials may round differently), Newton
    within 1e-8."""
    from pyoomph_amd.problem import Problem
    lam, c = 0.8, (2.0, 0.3)
    P = Problem(_face_pair(space, split, flux_form=_cubic_form(), flux=[(("right", "top"), c)], lam=lam))
    op = _cubic_oracle(space, split, lam, c)
    assert P.ndof == op.ndof
    x = np.random.default_rng(3).uniform(-0.3, 0.3, P.ndof)
    P.set_current_dofs(x)
    R, J = P.assemble_jacobian()
    r, rs, col, val = op.get_jacobian(x)
    assert np.array_equal(J.indptr, rs) and np.array_equal(J.indices, col)
    assert rel(J.data, val) <= 1e-12 and rel(R, r) <= 1e-12
    assert rel(P.get_residuals(dparameter="lambda"), op.get_residuals(x, deriv=1)) <= 1e-12
    _, dJ = P.assemble_jacobian(dparameter="lambda")
    _, drs, dcol, dval = op.get_jacobian(x, deriv=1)
    assert np.array_equal(dJ.indptr, drs) and np.array_equal(dJ.indices, dcol) and rel(dJ.data, dval) <= 1e-12
    # Newton from u = 0 where it converges (a milder flux and lambda)
    lam, c = 0.2, (0.5, 0.3)
    P = Problem(_face_pair(space, split, flux_form=_cubic_form(), flux=[(("right", "top"), c)], lam=lam))
    op = _cubic_oracle(space, split, lam, c)
    P.krylov, P.preconditioner = "gmres", "ilu0"
    st = P.solve()
    assert st[0]["converged"]
    assert rel(P.get_current_dofs(), on.newton_solve(op, np.zeros(op.ndof), tol=1e-12)) <= 1e-8


# ---- generated Hessian-vector products ------------------------------------------------------
@pytest.mark.parametrize("kind,kw", CASES)
def test_generated_hessian_products_match_the_compiled_ones(gpu, kind, kw):
    """sum_j y_j d2R_i/du_j du_k from the generated kernels (the Jacobian of the form's
    derivative in the direction y) against the compiled HV kernels: the same pattern, values
    within 1e-13 (the products may associate differently), plain and transposed alike."""
    from pyoomph_amd.problem import Problem
    a, b = pair(kind, **kw)
    A, B = Problem(a), Problem(b)
    rng = np.random.default_rng(6)
    x, y = rng.uniform(-0.3, 0.3, A.ndof), rng.uniform(-1, 1, A.ndof)
    A.set_current_dofs(x)
    B.set_current_dofs(x)
    for tr in (False, True):
        HA = A.assemble_hessian_vector_product(y, transposed=tr)
        HB = B.assemble_hessian_vector_product(y, transposed=tr)
        assert np.array_equal(HA.indptr, HB.indptr) and np.array_equal(HA.indices, HB.indices)
        if HA.nnz:
            assert rel(HB.data, HA.data) <= 1e-13


def test_generated_hessian_products_of_a_non_symmetric_form(gpu):
    """Nonlinear diffusion (1 + k u^2) grad u: H v = d/de J(x + e v) y and the transposed
    product d/de J(x + e v)^T y, by central differences of the generated Jacobian; the two
    products differ (the Jacobian is not symmetric)."""
    from pyoomph_amd.problem import Problem
    mesh = RectangularQuadMesh(6, split_in_tris="left", space="C2")
    from pyoomph_amd.problem import MeshProblem
    mp = MeshProblem(mesh, GeneratedEquation(library.nonlinear_diffusion(), space="C2", params=[2.0, 8.0]),
                     dirichlet=[(ALL2D, {0: 0.0})])
    P = Problem(mp)
    rng = np.random.default_rng(8)
    x, y, v = rng.uniform(-0.5, 0.5, P.ndof), rng.uniform(-1, 1, P.ndof), rng.uniform(-1, 1, P.ndof)
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
    assert np.max(np.abs((H - Ht).toarray())) > 1e-3 * np.max(np.abs(H.toarray()))


@pytest.mark.parametrize("space,split", FACE_CASES)
def test_generated_flux_with_the_normal_matches_the_oracle(gpu, space, split):
    """NeumannBC(u = c0 n_x + c1 n_y u) on the right, top and bottom sides (the normal from
    the tangent and each face's side): assembly within 1e-13 of the oracle's expression faces
    with its own outward normal, and on the right side alone the same residual as the
    constant flux c0 (n = (1, 0) there)."""
    from oracle import assembly as oa
    from oracle import faces as of
    from oracle import meshes as om
    from pyoomph_amd.codegen.face import FaceFluxForm
    from pyoomph_amd.problem import Problem
    c = (0.7, -1.3)
    form = FaceFluxForm("gen_normal_flux", lambda s: s.c[0] * s.n[0] + s.c[1] * s.n[1] * s.u, ncoef=2)
    sides = ("right", "top", "bottom")
    P = Problem(_face_pair(space, split, flux_form=form, flux=[(sides, c)]))
    m = om.rectangular_quad_mesh(7, split, space)
    faces = of.boundary_faces(m, list(sides))
    op = oa.OracleProblem(m, oa.FORM_BRATU, 1, m.on_boundary(["left"])[:, None], np.zeros((m.nnode, 1)), [1.0],
                          faces=faces, face_coef=np.tile(np.asarray(c), (len(faces), 1)), face_normal=True,
                          face_expr=(lambda u, x0, x1, C, n0, n1: C[:, 0] * n0 + (C[:, 1] * n1) * u,
                                     lambda u, x0, x1, C, n0, n1: C[:, 1] * n1, None, None))
    assert P.ndof == op.ndof
    x = np.random.default_rng(5).uniform(-0.3, 0.3, P.ndof)
    P.set_current_dofs(x)
    R, J = P.assemble_jacobian()
    r, rs, col, val = op.get_jacobian(x)
    assert np.array_equal(J.indptr, rs) and np.array_equal(J.indices, col)
    assert rel(J.data, val) <= 1e-13 and rel(R, r) <= 1e-13
    right = Problem(_face_pair(space, split, flux_form=form, flux=[(("right",), c)]))
    const = Problem(_face_pair(space, split, flux=[(("right",), 0.0, c[0])]))
    for Q in (right, const):
        Q.set_current_dofs(x)
    assert rel(right.get_residuals(), const.get_residuals()) <= 1e-14


def test_generated_taylor_hood_hessian_products(gpu):
    """Generated Navier-Stokes: the plain and transposed Hessian-vector products equal the
    hand-written kernel's (the convection block with u -> y) within 1e-12, and the plain one
    is the central difference of the Jacobian in the direction v."""
    from pyoomph_amd.problem import Problem
    mp_a, mp_b = _th_pair(6, 80.0)
    A, B = Problem(mp_a), Problem(mp_b)
    rng = np.random.default_rng(9)
    x, y, v = (rng.uniform(-0.5, 0.5, A.ndof) for _ in range(3))
    A.set_current_dofs(x)
    B.set_current_dofs(x)
    for tr in (False, True):
        HA = A.assemble_hessian_vector_product(y, transposed=tr)
        HB = B.assemble_hessian_vector_product(y, transposed=tr)
        assert np.array_equal(HA.indptr, HB.indptr) and np.array_equal(HA.indices, HB.indices)
        assert rel(HB.data, HA.data) <= 1e-12, tr
    H = B.assemble_hessian_vector_product(y)
    eps = 1e-5
    B.set_current_dofs(x + eps * v)
    Jp = B.assemble_jacobian()[1]
    B.set_current_dofs(x - eps * v)
    Jm = B.assemble_jacobian()[1]
    fd = (Jp @ y - Jm @ y) / (2 * eps)
    assert np.max(np.abs(H @ v - fd)) <= 1e-7 * np.max(np.abs(fd))
