"""Time stepping on the GPU against the oracle: the time derivative of the scalar forms and
of the Taylor-Hood momentum with the reference's BDF weights, the mass matrix (flag 2),
unsteady Newton steps.

Bar: as the steady parity (CSR pattern bit-exact, FP64 values bitwise except the Bratu
exponential, ≤1e-12), solutions of every time step within 1e-8 of the oracle's SuperLU
Newton, and a sub-problem's steps give the same bits alone and in a batch."""
import numpy as np
import pytest

from oracle import timestepping as ots
from tests.oracle_bridge import oracle_problem, product_problem
from tests.test_gpu_parity import build, rel

pytestmark = pytest.mark.gpu

CASES = [("poisson", dict(N=5)), ("bratu", dict(N=7, space="C1", lam=1.5)),
         ("bratu", dict(N=5, space="C2", lam=2.5)), ("rd", dict(N=3)), ("ns", dict(N=4, Re=50.0))]


@pytest.mark.parametrize("kind,kw", CASES)
def test_transient_assembly_parity(gpu, kind, kw):
    import torch
    P, op = build(kind, kw)
    rng = np.random.default_rng(7)
    x, h1, h2 = (rng.uniform(-0.4, 0.4, P.ndof) for _ in range(3))
    tw = ots.bdf2_weights(0.05, 0.08)
    P.set_current_dofs(x)
    d1 = torch.tensor(h1, dtype=torch.float64, device=P.device)
    d2 = torch.tensor(h2, dtype=torch.float64, device=P.device)
    P.set_time_state(tw, d1, d2)
    try:
        R, J = P.assemble_jacobian()
    finally:
        P.set_time_state(None, None, None)
    op.set_time(tw, h1, h2)
    r, rs, col, val = op.get_jacobian(x)
    op.set_time(None, None, None)
    assert np.array_equal(J.indptr, rs) and np.array_equal(J.indices, col)
    assert rel(J.data, val) <= 1e-12 and rel(R, r) <= 1e-12
    if kind != "bratu":
        assert np.array_equal(J.data, val) and np.array_equal(R, r)


@pytest.mark.parametrize("kind,kw", CASES[:2] + CASES[3:])
def test_mass_matrix_parity(gpu, kind, kw):
    P, op = build(kind, kw)
    x = np.random.default_rng(3).uniform(-0.4, 0.4, P.ndof)
    P.set_current_dofs(x)
    M = P.assemble_mass_matrix()
    op.set_time(ots.bdf1_weights(1.0), x, x)
    _, rs, col, val = op.get_mass_matrix(x)
    op.set_time(None, None, None)
    assert np.array_equal(M.indptr, rs) and np.array_equal(M.indices, col)
    assert np.array_equal(M.data, val)


@pytest.mark.parametrize("kind,kw,opts", [
    ("bratu", dict(N=10, space="C2", lam=2.0), dict(krylov="cg", preconditioner="jacobi")),
    ("bratu", dict(N=12, space="C1", lam=3.0), dict(krylov="cg", preconditioner="amg")),
    ("poisson", dict(N=8), dict(krylov="gmres", preconditioner="ilu0")),
    ("ns", dict(N=6, Re=100.0), dict(krylov="gmres", preconditioner="ilu0")),
])
def test_unsteady_steps_match_oracle(gpu, kind, kw, opts):
    P, op = build(kind, kw)
    for k, v in opts.items():
        setattr(P, k, v)
    if kind == "ns":
        P.newton_solver_tolerance = 1e-10
    P.set_initial_condition(np.zeros(P.ndof))
    st = ots.TimeState(np.zeros(op.ndof))
    xo = np.zeros(op.ndof)
    for dt in (0.05, 0.05, 0.08, 0.03):
        s = P.unsteady_newton_solve(dt)
        assert s[0]["converged"]
        xo = ots.unsteady_newton_step(op, xo, st, dt, tol=P.newton_solver_tolerance)
        assert rel(P.get_current_dofs(), xo) <= 1e-8
    assert P.time_steps_done == 4 and abs(P.time - st.time) < 1e-15


def test_transient_batch_equals_alone_bitwise(gpu):
    from pyoomph_amd.problem import Problem
    specs = [("bratu", dict(N=9, space="C2", lam=2.0)), ("poisson", dict(N=6)),
             ("bratu", dict(N=14, space="C1", lam=1.0))]

    def run(sp):
        P = Problem([product_problem(k, **kw) for k, kw in sp])
        P.krylov, P.preconditioner = "cg", "jacobi"
        P.set_initial_condition(np.zeros(P.ndof))
        for dt in (0.1, 0.1, 0.05):
            P.unsteady_newton_solve(dt)
        return P, P.get_current_dofs()

    Pb, xb = run(specs)
    for i, sp in enumerate(specs):
        Pa, xa = run([sp])
        assert np.array_equal(Pb.sub_dofs(i, xb), Pa.sub_dofs(0, xa)), i
