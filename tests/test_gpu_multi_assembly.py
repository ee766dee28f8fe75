"""GPU multi-assembly (the reference's Problem::assemble_multiassembly,
src/problem.cpp:2326-2368, through MultiAssembleRequest,
pyoomph/generic/bifurcation_tools.py:442-520) against the oracle.

Every matrix comes in the multi-assembly's CSR form: the oracle's vectors-of-pairs matrix
(exact-zero element contributions dropped, element-order sums) with each row's columns in
ascending order (map_order_csr). Bar: the pattern bit-exact; values and vectors within
1e-12 relative, bitwise where no exponential enters (everything but Bratu)."""
import numpy as np
import pytest

from oracle import timestepping as ots
from oracle.assembly import map_order_csr
from tests.oracle_bridge import oracle_problem, product_problem

pytestmark = pytest.mark.gpu


def build(kind, kw):
    from pyoomph_amd.problem import Problem
    mp = product_problem(kind, **kw)
    okw = dict(kw)
    if kind == "rd":
        okw["f"] = mp.equations.source_field
    op = oracle_problem(kind, **okw)
    assert np.array_equal(mp.eqn, op.eqn)
    return Problem(mp), op


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def same_matrix(M, oracle_csr, exact):
    rs, col, val = map_order_csr(*oracle_csr)
    assert M.has_sorted_indices
    assert np.array_equal(M.indptr, rs) and np.array_equal(M.indices, col)
    if val.size == 0:
        assert M.nnz == 0
    elif exact:
        assert np.array_equal(M.data, val)
    else:
        assert rel(M.data, val) <= 1e-12


def oracle_mass(op, x):
    op.set_time(ots.bdf1_weights(1.0), x, x)
    try:
        return op.get_mass_matrix(x)[1:]
    finally:
        op.set_time(None, None, None)


@pytest.mark.parametrize("kind,kw", [("bratu", dict(N=6, space="C1", lam=1.5)),
                                     ("bratu", dict(N=5, space="C2", lam=2.5))])
def test_bratu_every_quantity(gpu, kind, kw):
    """R, J, M, dR/dp, dJ/dp, dM/dp, H y and M_H y in one call, in request order."""
    P, op = build(kind, kw)
    rng = np.random.default_rng(4)
    x, y = rng.uniform(0.0, 0.4, P.ndof), rng.uniform(-1, 1, P.ndof)
    P.set_current_dofs(x)
    out = (P.start_multiassembly().dJdU(y).R().J().dRdp("lambda").M().dJdp("lambda")
           .dMdp("lambda").dMdU(y).assemble())
    H, R, J, dR, M, dJ, dM, MH = out
    assert rel(R, op.get_residuals(x)) <= 1e-12
    assert rel(dR, op.get_residuals(x, deriv=1)) <= 1e-12
    same_matrix(J, op.get_jacobian(x)[1:], exact=False)
    same_matrix(dJ, op.get_jacobian(x, deriv=1)[1:], exact=False)
    same_matrix(M, oracle_mass(op, x), exact=True)
    same_matrix(H, op.get_hessian_vector_product(x, y)[1:], exact=False)
    # the mass coefficient depends neither on lambda nor on u: no entry survives
    assert dM.nnz == 0 and MH.nnz == 0 and dM.shape == (P.ndof, P.ndof)
    # the same quantities from the single-quantity entry points, as scipy matrices
    R1, J1 = P.assemble_jacobian()
    assert np.array_equal(R, R1) and (J - J1).count_nonzero() == 0
    assert (J.sorted_indices().data == J1.sorted_indices().data).all()


@pytest.mark.parametrize("transposed", [False, True])
def test_navier_stokes_products(gpu, transposed):
    """Taylor-Hood: plain and transposed Hessian products of two vectors (the transposed
    one is the Hessian of y . R) with R, J and M, against the oracle bit for bit."""
    P, op = build("ns", dict(N=5, Re=60.0))
    rng = np.random.default_rng(8)
    x = rng.uniform(-0.4, 0.4, P.ndof)
    y0, y1 = rng.uniform(-1, 1, P.ndof), rng.uniform(-1, 1, P.ndof)
    P.set_current_dofs(x)
    R, J, M, H0, H1, MH = (P.start_multiassembly().R().J().M().dJdU(y0, transposed=transposed)
                           .dJdU(y1, transposed=transposed).dMdU(y0, transposed=transposed).assemble())
    assert np.array_equal(R, op.get_residuals(x))
    same_matrix(J, op.get_jacobian(x)[1:], exact=True)
    same_matrix(M, oracle_mass(op, x), exact=True)
    same_matrix(H0, op.get_hessian_vector_product(x, y0, transposed)[1:], exact=True)
    same_matrix(H1, op.get_hessian_vector_product(x, y1, transposed)[1:], exact=True)
    assert MH.nnz == 0
    if transposed:
        assert abs(H0 - H0.T).max() <= 1e-13 * abs(H0).max()
        # the single-vector entry point gives the same matrix
        assert (P.assemble_hessian_vector_product(y0, transposed=True) - H0).count_nonzero() == 0


def test_reaction_diffusion_and_transient_state(gpu):
    """Hex Q2 reaction-diffusion with time stepping on: the mass matrix does not depend on
    the BDF state, the residual carries the time term (as a plain assembly does)."""
    P, op = build("rd", dict(N=3))
    rng = np.random.default_rng(6)
    x, y = rng.uniform(-0.4, 0.4, P.ndof), rng.uniform(-1, 1, P.ndof)
    h1, h2 = rng.uniform(-0.4, 0.4, P.ndof), rng.uniform(-0.4, 0.4, P.ndof)
    P.set_current_dofs(x)
    M0, H0 = P.start_multiassembly().M().dJdU(y, transposed=True).assemble()
    import torch
    w = P.bdf_weights(0.1, 0.2)
    d1, d2 = (torch.tensor(h, device=P.device) for h in (h1, h2))
    P.set_time_state(w, d1, d2)
    op.set_time(w, h1, h2)
    try:
        R, M1 = P.start_multiassembly().R().M().assemble()
        assert np.array_equal(R, op.get_residuals(x))
        same_matrix(M1, op.get_mass_matrix(x)[1:], exact=True)
    finally:
        P.set_time_state(None, None, None)
        op.set_time(None, None, None)
    assert (M0 - M1).count_nonzero() == 0
    same_matrix(H0, op.get_hessian_vector_product(x, y)[1:], exact=True)


def test_faces_contribute_to_residual_and_jacobian_only(gpu):
    """Robin faces (linear in u, no time term): in R and J, not in M or H."""
    from tests.test_gpu_faces import build as build_faces
    P, op = build_faces(6, "C2", 1.2, 2.0, 0.5)
    rng = np.random.default_rng(2)
    x, y = rng.uniform(-0.3, 0.3, P.ndof), rng.uniform(-1, 1, P.ndof)
    P.set_current_dofs(x)
    R, J, H, M = P.start_multiassembly().R().J().dJdU(y).M().assemble()
    assert rel(R, op.get_residuals(x)) <= 1e-12
    same_matrix(J, op.get_jacobian(x)[1:], exact=False)
    same_matrix(H, op.get_hessian_vector_product(x, y)[1:], exact=False)
    same_matrix(M, oracle_mass(op, x), exact=True)


def test_requests_are_checked_like_the_reference(gpu):
    P, _ = build("bratu", dict(N=4, space="C1", lam=1.0))
    y = np.ones(P.ndof)
    with pytest.raises(RuntimeError, match="multiple 'jacobian'"):
        P.start_multiassembly().J().R().J().assemble()
    with pytest.raises(RuntimeError, match="transposed and non-transposed"):
        P.start_multiassembly().dJdU(y).dJdU(y, transposed=True).assemble()
    with pytest.raises(RuntimeError, match="contribution"):
        P.start_multiassembly().R("interface").assemble()
    with pytest.raises(ValueError, match="unknown global parameter"):
        P.start_multiassembly().dRdp("mu").assemble()
    # the problem stays usable
    R, = P.start_multiassembly().R().assemble()
    assert np.array_equal(R, P.get_residuals())


def test_taylor_hood_has_no_parameter_derivative(gpu):
    P, _ = build("ns", dict(N=3))
    with pytest.raises((RuntimeError, ValueError)):
        P.start_multiassembly().dRdp("lambda").assemble()


MASS = "_simple_mass_matrix_of_defined_fields"


def _bratu_with_mass_table(N, space, lam):
    """A generated Bratu form with a second residual table int u psi_i (the mass matrix of
    the field, the table pyoomph's stability setup adds with destination=...,
    pyoomph/generic/problem.py:1677-1684)."""
    from pyoomph_amd.codegen import expr as E
    from pyoomph_amd.codegen.runtime import GeneratedEquation
    from pyoomph_amd.codegen.scalar import ScalarWeakForm
    from pyoomph_amd.meshes import RectangularQuadMesh
    from pyoomph_amd.problem import MeshProblem
    form = ScalarWeakForm("gen_bratu_tables", lambda s: [s.grad_u[0], s.grad_u[1]],
                          lambda s: -s.p[0] * E.exp(s.u), nparams=1, global_param=0,
                          tables={MASS: (lambda s: [0, 0], lambda s: s.u)})
    eq = GeneratedEquation(form, space=space, params=[lam])
    return MeshProblem(RectangularQuadMesh(N, split_in_tris="left", space=space), eq,
                       dirichlet=[(("left", "right", "top", "bottom"), {0: 0.0})])


@pytest.mark.parametrize("space", ["C1", "C2"])
def test_residual_tables(gpu, space):
    """A named residual table: assemble_jacobian(which_one=...) gives the oracle's mass matrix
    bit for bit and restores the solved residual; multi-assembly contributions; a batch with
    a compiled form (no such table) contributes nothing to it; unknown names raise."""
    from pyoomph_amd.problem import Problem
    mp = _bratu_with_mass_table(6, space, 1.3)
    op = oracle_problem("bratu", N=6, space=space, lam=1.3)
    assert np.array_equal(mp.eqn, op.eqn)
    P = Problem(mp)
    x = np.random.default_rng(2).uniform(-0.3, 0.3, P.ndof)
    P.set_current_dofs(x)
    R0, J0 = P.assemble_jacobian()
    Rm, Jm = P.assemble_jacobian(which_one=MASS)
    op.set_time(ots.bdf1_weights(1.0), x, x)
    _, rs, col, val = op.get_mass_matrix(x)
    op.set_time(None, None, None)
    assert np.array_equal(Jm.indptr, rs) and np.array_equal(Jm.indices, col) and np.array_equal(Jm.data, val)
    assert np.allclose(Rm, Jm @ x, rtol=1e-12, atol=1e-15)     # int u psi_i = M u
    assert P._get_solved_residual() == ""
    R1, J1 = P.assemble_jacobian()
    assert np.array_equal(R0, R1) and np.array_equal(J0.data, J1.data)
    # multi-assembly over two contributions
    Jd, Jt, Rt = P.start_multiassembly().J().J(MASS).R(MASS).assemble()
    assert (Jd - J0).count_nonzero() == 0 and (Jt - Jm).count_nonzero() == 0 and np.array_equal(Rt, Rm)
    with pytest.raises(RuntimeError, match="no equations at all"):
        P._set_solved_residual("nothing")
    assert P._set_solved_residual("nothing", raise_error=False) is False
    assert P._get_solved_residual() == ""
    # batched with a compiled Bratu problem: its rows have no table, so no entries
    from pyoomph_amd import workloads as wl
    B = Problem([mp, wl.bratu(5, "C1", 1.0)])
    xb = np.random.default_rng(3).uniform(-0.3, 0.3, B.ndof)
    B.set_current_dofs(xb)
    Rb, Jb = B.assemble_jacobian(which_one=MASS)
    o = B.dof_offsets
    assert np.all(Rb[o[1]:] == 0.0) and Jb[o[1]:].nnz == 0
    P.set_current_dofs(xb[:o[1]])
    Ra, Ja = P.assemble_jacobian(which_one=MASS)
    assert np.array_equal(Rb[:o[1]], Ra) and (Jb[:o[1], :o[1]] - Ja).count_nonzero() == 0
    st = P.solve()
    assert st[0]["converged"]


def test_generated_taylor_hood_mass_derivatives_are_refused(gpu):
    """A generated Taylor-Hood form's mass coefficient may depend on the parameter and the
    dofs: dM/dp and the mass-matrix Hessian product are refused, the Hessian product of
    the Jacobian (generated kernels) is not."""
    from pyoomph_amd.problem import Problem
    from tests.test_gpu_codegen import _th_pair
    _, mp = _th_pair(4, 50.0)
    P = Problem(mp)
    y = np.random.default_rng(2).uniform(-1, 1, P.ndof)
    P.set_current_dofs(0.1 * y)
    H, = P.start_multiassembly().dJdU(y).assemble()
    assert H.nnz > 0
    with pytest.raises(RuntimeError, match="mass-matrix Hessian"):
        P.start_multiassembly().dMdU(y).assemble()


def test_generated_taylor_hood_mass_parameter_derivative(gpu):
    """A generated Taylor-Hood form whose mass coefficient is the global parameter squared:
    dM/dp from the multi-assembly equals the central difference of its mass matrix in the
    parameter."""
    from pyoomph_amd.codegen import library
    from pyoomph_amd.codegen.mixed import TaylorHoodWeakForm
    from pyoomph_amd.codegen.runtime import GeneratedEquation
    from pyoomph_amd.meshes import RectangularQuadMesh
    from pyoomph_amd.problem import MeshProblem, Problem
    ns = library.navier_stokes()
    form = TaylorHoodWeakForm("gen_th_pmass", ns.flux_fn, ns.source_fn, ns.cont_fn,
                              mass=lambda s: s.par[2] * s.par[2], nparams=3, global_param=2)
    mesh = RectangularQuadMesh(4, split_in_tris=False, space="C2")
    bcs = [(("top",), {0: 1.0, 1: 0.0}), (("left", "right", "bottom"), {0: 0.0, 1: 0.0})]

    def problem(p):
        eq = GeneratedEquation(form, space="C2", params=[1.0 / 50.0, 1.0, p])
        return Problem(MeshProblem(mesh, eq, dirichlet=bcs, pin_extra=[(0, 2, 0.0)]))
    p0, h = 1.3, 1e-4
    P = problem(p0)
    x = np.random.default_rng(5).uniform(-0.3, 0.3, P.ndof)
    P.set_current_dofs(x)
    dM, = P.start_multiassembly().dMdp("lambda").assemble()
    Ms = []
    for p in (p0 + h, p0 - h):
        Q = problem(p)
        Q.set_current_dofs(x)
        Ms.append(Q.start_multiassembly().M().assemble()[0].toarray())
    fd = (Ms[0] - Ms[1]) / (2 * h)
    assert dM.nnz > 0
    assert np.max(np.abs(dM.toarray() - fd)) <= 1e-8 * np.max(np.abs(fd))
