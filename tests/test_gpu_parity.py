"""GPU parity: the HIP path against the oracle on the same seeded inputs.

Bar: CSR pattern (row starts, columns in first-touch order, exact-zero drop) bit-exact;
FP64 values and residuals within 1e-12 relative (bitwise where no transcendental
function enters, i.e. everything but the Bratu exponential); Newton solutions within
1e-8 relative of the oracle's SuperLU Newton.
"""
import numpy as np
import pytest

from oracle import newton as on
from tests.oracle_bridge import oracle_problem, product_problem

pytestmark = pytest.mark.gpu

CASES = [("poisson", dict(N=8)), ("poisson", dict(N=5, seed=3)),
         ("bratu", dict(N=9, space="C1", lam=1.5)), ("bratu", dict(N=7, space="C2", lam=2.5)),
         ("rd", dict(N=3)), ("ns", dict(N=6)), ("ns", dict(N=3, Re=10.0))]


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


@pytest.mark.parametrize("kind,kw", CASES)
def test_assembly_parity(gpu, kind, kw):
    P, op = build(kind, kw)
    for seed in (1, 2):
        x = np.random.default_rng(seed).uniform(-0.4, 0.4, P.ndof)
        P.set_current_dofs(x)
        R, J = P.assemble_jacobian()
        r, rs, col, val = op.get_jacobian(x)
        assert np.array_equal(J.indptr, rs)
        assert np.array_equal(J.indices, col)
        assert rel(J.data, val) <= 1e-12
        assert rel(R, r) <= 1e-12
        if kind != "bratu":
            # no transcendental function in the Jacobians: bitwise
            assert np.array_equal(J.data, val)
        if kind not in ("bratu", "poisson"):
            # the Poisson residual holds exp of the Gaussian sources, where device exp and
            # glibc exp may round differently; everything else is bitwise
            assert np.array_equal(R, r)
        assert np.array_equal(P.get_residuals(), R)


def test_zero_state_pattern_drops_exact_zeros(gpu):
    """NS at the zero state: the convective blocks vanish; exact zeros are dropped
    per contribution exactly as the oracle (vectors_of_pairs) drops them."""
    P, op = build("ns", dict(N=4))
    P.set_current_dofs(np.zeros(P.ndof))
    R, J = P.assemble_jacobian()
    r, rs, col, val = op.get_jacobian(np.zeros(op.ndof))
    assert np.array_equal(J.indptr, rs) and np.array_equal(J.indices, col)
    assert np.array_equal(J.data, val)
    assert J.nnz < P.nnz     # structural pattern is a strict superset


def test_bratu_parameter_derivative(gpu):
    P, op = build("bratu", dict(N=6, space="C2", lam=1.7))
    x = np.random.default_rng(7).uniform(0, 0.3, P.ndof)
    P.set_current_dofs(x)
    d = P.get_residuals(dparameter="lambda")
    assert rel(d, op.get_residuals(x, deriv=1)) <= 1e-12


def test_assembly_is_deterministic(gpu):
    P, _ = build("bratu", dict(N=20, space="C2", lam=2.0))
    P.set_current_dofs(np.random.default_rng(0).uniform(0, 0.2, P.ndof))
    a = [P.assemble_device("f64") for _ in range(3)]
    for r, v in a[1:]:
        assert np.array_equal(r.cpu().numpy(), a[0][0].cpu().numpy())
        assert np.array_equal(v.cpu().numpy(), a[0][1].cpu().numpy())


@pytest.mark.parametrize("kind,kw", [("bratu", dict(N=11, space="C2", lam=2.0)),
                                     ("ns", dict(N=5)), ("rd", dict(N=3))])
def test_fp32_kernels_close_to_fp64(gpu, kind, kw):
    P, _ = build(kind, kw)
    P.set_current_dofs(np.random.default_rng(1).uniform(-0.3, 0.3, P.ndof))
    r64, v64 = P.assemble_device("f64")
    r32, v32 = P.assemble_device("f32")
    v64, v32 = v64.cpu().numpy(), v32.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(v32 - v64)) <= 2e-6 * np.max(np.abs(v64))
    assert np.max(np.abs(r32.cpu().numpy() - r64.cpu().numpy())) <= 2e-6 * np.max(np.abs(v64))


@pytest.mark.parametrize("kind,kw,opts", [
    ("poisson", dict(N=16), dict(krylov="cg", preconditioner="jacobi")),
    ("poisson", dict(N=12), dict(krylov="gmres", preconditioner="ilu0")),
    ("bratu", dict(N=20, space="C2", lam=2.0), dict(krylov="cg", preconditioner="ilu0")),
    ("bratu", dict(N=17, space="C1", lam=3.0), dict(krylov="cg", preconditioner="jacobi",
                                                   linear_precision="f32")),
    ("bratu", dict(N=10, space="C2", lam=1.0), dict(krylov="cg", preconditioner="jacobi",
                                                   linear_precision="f32", jacobian_precision="f32")),
    ("rd", dict(N=3), dict(krylov="cg", preconditioner="ilu0")),
    ("ns", dict(N=8), dict(krylov="gmres", preconditioner="ilu0")),
    ("ns", dict(N=12), dict(krylov="gmres", preconditioner="ilu0", ilu_order="natural",
                            gmres_restart=100)),
    ("ns", dict(N=16), dict(krylov="gmres", preconditioner="ilu0", ilu_order="multicolor",
                            gmres_restart=150)),
    ("poisson", dict(N=20), dict(krylov="gmres", preconditioner="ilu0", ilu_order="multicolor")),
    ("poisson", dict(N=24, seed=2), dict(krylov="cg", preconditioner="amg")),
    ("bratu", dict(N=40, space="C2", lam=2.5), dict(krylov="cg", preconditioner="amg")),
    ("bratu", dict(N=33, space="C1", lam=0.7), dict(krylov="cg", preconditioner="amg")),
    ("rd", dict(N=6), dict(krylov="cg", preconditioner="amg")),
    ("bratu", dict(N=12, space="C2", lam=1.5), dict(krylov="gmres", preconditioner="amg")),
])
def test_newton_matches_superlu_oracle(gpu, kind, kw, opts):
    P, op = build(kind, kw)
    for k, v in opts.items():
        setattr(P, k, v)
    if kind == "ns":
        P.newton_solver_tolerance = 1e-10
    st = P.solve()
    assert st[0]["converged"] and st[0]["linear_failures"] == 0
    xo = on.newton_solve(op, np.zeros(op.ndof), tol=P.newton_solver_tolerance)
    assert rel(P.get_current_dofs(), xo) <= 1e-8


def test_batch_equals_alone_bitwise(gpu):
    """Per-sub-problem reductions: a problem gives the same bits alone and in a batch."""
    from pyoomph_amd import workloads as wl
    from pyoomph_amd.problem import Problem
    specs = [(13, "C2", 1.2), (20, "C1", 2.9), (9, "C2", 0.7), (16, "C1", 1.9)]
    batch = Problem([wl.bratu(*s) for s in specs])
    batch.krylov, batch.preconditioner = "cg", "jacobi"
    batch.solve()
    xb = batch.get_current_dofs()
    for i, s in enumerate(specs):
        alone = Problem(wl.bratu(*s))
        alone.krylov, alone.preconditioner = "cg", "jacobi"
        alone.solve()
        assert np.array_equal(alone.get_current_dofs(), batch.sub_dofs(i, xb))


@pytest.mark.parametrize("kind,kw,opts", [
    ("bratu", dict(N=12, space="C2", lam=2.0), dict(krylov="cg", preconditioner="jacobi")),
    ("rd", dict(N=3), dict(krylov="cg", preconditioner="jacobi")),
    ("ns", dict(N=5), dict(krylov="gmres", preconditioner="ilu0")),
    ("ns", dict(N=9), dict(krylov="gmres", preconditioner="ilu0", gmres_restart=100)),
    ("bratu", dict(N=40, space="C2", lam=2.0), dict(krylov="cg", preconditioner="amg")),
    ("rd", dict(N=6), dict(krylov="cg", preconditioner="amg")),
])
def test_spmv_lane_count_keeps_bits(gpu, monkeypatch, kind, kw, opts):
    """Every spmv variant (1, 2, 4, 8 lanes per row; LDS staging in chunks of 512, 1024,
    2048 or 4096 entries, with rows that span chunks) sums a row in the same order."""
    xs = []
    for lpr in ("1", "2", "4", "8", "16", "32", "64", "128"):
        monkeypatch.setenv("PYHIP_SPMV_LPR", lpr)
        P, op = build(kind, kw)
        for k, v in opts.items():
            setattr(P, k, v)
        if kind == "ns":
            P.newton_solver_tolerance = 1e-10
        st = P.solve()
        assert st[0]["converged"]
        xs.append(P.get_current_dofs())
        del P
    for x in xs[1:]:
        assert np.array_equal(x, xs[0])
    xo = on.newton_solve(op, np.zeros(op.ndof), tol=1e-10 if kind == "ns" else 1e-8)
    assert rel(xs[0], xo) <= 1e-8


def test_cg_iteration_cap_reports_failure(gpu):
    """A CG capped below what a problem needs reports the failure instead of a silent step."""
    P, _ = build("bratu", dict(N=24, space="C2", lam=1.0))
    P.krylov, P.preconditioner = "cg", "jacobi"
    P.gmres_max_iters = 3
    st = P.solve(raise_on_failure=False)
    assert st[0]["linear_failures"] > 0


def test_multigrid_batch_equals_alone_bitwise(gpu):
    """The multigrid hierarchy is built per sub-problem: the same bits alone and in a batch."""
    from pyoomph_amd import workloads as wl
    from pyoomph_amd.problem import Problem
    specs = [(45, "C2", 1.2), (70, "C1", 2.9), (9, "C2", 0.7), (3, "C1", 1.0), (26, "C1", 1.9)]
    batch = Problem([wl.bratu(*s) for s in specs])
    batch.krylov, batch.preconditioner = "cg", "amg"
    st = batch.solve()
    xb = batch.get_current_dofs()
    for i, s in enumerate(specs):
        alone = Problem(wl.bratu(*s))
        alone.krylov, alone.preconditioner = "cg", "amg"
        sa = alone.solve()
        assert np.array_equal(alone.get_current_dofs(), batch.sub_dofs(i, xb)), s
        assert sa[0]["newton_iters"] == st[i]["newton_iters"]


def test_newton_error_path(gpu):
    from pyoomph_amd.problem import NewtonSolverError
    P, _ = build("bratu", dict(N=8, space="C2", lam=2.5))
    P.max_newton_iterations = 1
    P.newton_solver_tolerance = 1e-14
    with pytest.raises(NewtonSolverError):
        P.solve()
    st = P.solve(raise_on_failure=False)
    assert not st[0]["converged"]


def test_full_size_poisson_against_oracle(gpu):
    """64 x 64 P2 Poisson (16,129 dofs): CSR bit-exact and solution at 1e-8."""
    P, op = build("poisson", dict(N=64, seed=0))
    x = np.random.default_rng(11).uniform(-1, 1, P.ndof)
    P.set_current_dofs(x)
    R, J = P.assemble_jacobian()
    r, rs, col, val = op.get_jacobian(x)
    assert np.array_equal(J.indptr, rs) and np.array_equal(J.indices, col)
    assert rel(J.data, val) <= 1e-12 and rel(R, r) <= 1e-12
    P.set_current_dofs(np.zeros(P.ndof))
    P.krylov, P.preconditioner = "cg", "ilu0"
    P.solve()
    xo = on.newton_solve(op, np.zeros(op.ndof))
    assert rel(P.get_current_dofs(), xo) <= 1e-8


def test_ilu_order_is_decided_per_sub_problem(gpu):
    """With the automatic ILU(0) order, a small cavity (natural order, fewer than 64
    levels) keeps its order next to a large one (multicolour): same bits alone and in the
    batch."""
    from pyoomph_amd import workloads as wl
    from pyoomph_amd.problem import Problem

    def solve(parts):
        P = Problem(parts)
        P.krylov, P.preconditioner, P.gmres_restart = "gmres", "ilu0", 100
        P.newton_solver_tolerance = 1e-10
        st = P.solve()
        assert all(s["converged"] for s in st)
        return P, P.get_current_dofs()

    Pb, xb = solve([wl.lid_driven_cavity(24), wl.lid_driven_cavity(4)])
    Ps, xs = solve([wl.lid_driven_cavity(4)])
    assert np.array_equal(Pb.sub_dofs(1, xb), Ps.sub_dofs(0, xs))


def test_cavity_64_multicolor_ilu_against_oracle(gpu):
    """Lid-driven cavity at 64 x 64 Taylor-Hood elements (36 k unknowns) from u = 0 to
    max|R| < 1e-10: GMRES(300) with multicolour ILU(0) against the SuperLU Newton."""
    P, op = build("ns", dict(N=64))
    P.krylov, P.preconditioner, P.ilu_order, P.gmres_restart = "gmres", "ilu0", "auto", 300
    P.gmres_max_iters = 8000
    P.newton_solver_tolerance = 1e-10
    st = P.solve()
    assert st[0]["converged"] and st[0]["final_maxres"] < 1e-10
    xo = on.newton_solve(op, np.zeros(op.ndof), tol=1e-10)
    assert rel(P.get_current_dofs(), xo) <= 1e-8


def test_solver_plugin_matches_superlu(gpu):
    """pyhip_solve_serial with oomph's arguments (CSR as CSC of J^T, transpose=1)."""
    import ctypes
    from pyoomph_amd import _lib
    op = oracle_problem("ns", N=4)
    x = np.random.default_rng(5).uniform(-0.3, 0.3, op.ndof)
    r, rs, col, val = op.get_jacobian(x)
    n, nnz = op.ndof, len(val)
    ref = on.superlu_solve(n, rs, col, val, r)
    L = _lib.load()
    h = ctypes.c_void_p()
    v = np.ascontiguousarray(val)
    ri = np.ascontiguousarray(col, dtype=np.int32)
    cp = np.ascontiguousarray(rs, dtype=np.int32)
    _lib.check(L.pyhip_solve_serial(1, n, nnz, 1, v.ctypes.data_as(_lib.c_dp), ri.ctypes.data_as(_lib.c_i32p),
                                    cp.ctypes.data_as(_lib.c_i32p), None, n, 1, ctypes.byref(h)))
    b = np.ascontiguousarray(r.copy())
    _lib.check(L.pyhip_solve_serial(2, n, nnz, 1, None, None, None, b.ctypes.data_as(_lib.c_dp), n, 1,
                                    ctypes.byref(h)))
    _lib.check(L.pyhip_solve_serial(3, n, nnz, 1, None, None, None, None, n, 1, ctypes.byref(h)))
    assert rel(b, ref) <= 1e-8


@pytest.mark.parametrize("kind,kw", [("bratu", dict(N=7, space="C1", lam=1.5)),
                                     ("bratu", dict(N=5, space="C2", lam=2.5)),
                                     ("rd", dict(N=3)), ("poisson", dict(N=4)), ("ns", dict(N=4, Re=70.0))])
def test_hessian_vector_product_parity(gpu, kind, kw):
    """sum_k dJ/du_k y_k: reference CSR (first-touch, exact zeros dropped) and values
    against the oracle; Poisson's product is identically zero, so no entry survives;
    Taylor-Hood's is the convection block with u -> y."""
    P, op = build(kind, kw)
    rng = np.random.default_rng(11)
    x, y = rng.uniform(-0.4, 0.4, P.ndof), rng.uniform(-1, 1, P.ndof)
    P.set_current_dofs(x)
    H = P.assemble_hessian_vector_product(y)
    r, rs, col, val = op.get_hessian_vector_product(x, y)
    assert np.array_equal(H.indptr, rs) and np.array_equal(H.indices, col)
    assert rel(H.data, val) <= 1e-12 if val.size else H.nnz == 0
    if kind != "bratu":
        assert np.array_equal(H.data, val)


def test_hessian_products_of_several_vectors(gpu):
    """Y [nvec, ndof] gives one product per vector, each equal to its single-vector call."""
    P, op = build("ns", dict(N=3, Re=40.0))
    rng = np.random.default_rng(13)
    P.set_current_dofs(rng.uniform(-0.4, 0.4, P.ndof))
    Y = rng.uniform(-1, 1, (3, P.ndof))
    Hs = P.assemble_hessian_vector_product(Y)
    assert len(Hs) == 3
    for k in range(3):
        H1 = P.assemble_hessian_vector_product(Y[k])
        assert np.array_equal(Hs[k].indptr, H1.indptr) and np.array_equal(Hs[k].data, H1.data)
    # linear in the vector: H(y0 + 2 y1) = H(y0) + 2 H(y1) on the structural entries
    Hl = P.assemble_hessian_vector_product(Y[0] + 2.0 * Y[1])
    D = (Hl - (Hs[0] + 2.0 * Hs[1])).toarray()
    assert np.max(np.abs(D)) <= 1e-12 * np.max(np.abs(Hl.toarray()))
