"""The JIT element table's per-element slots and the superlu() entry on the GPU.

Each slot is called the way oomph's element loop calls generated code: one element's
JITElementInfo_t (nodal values, coordinates, local equations with -1 for pinned values),
a shape info whose jacobian_size is the stride of J, caller-zeroed R / J / M buffers. The
results must equal the oracle's element outputs in the element's local numbering: bitwise
(the same device kernel as the batched assembly, whose FP64 values equal the oracle's
bitwise), within 1e-12 where exp enters (Bratu, the Poisson source). The superlu() entry
factorises the oracle's CSR Jacobian as oomph hands it over (compressed columns of J^T,
transpose set) and solves J x = b to 1e-10 of scipy's splu.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from pyoomph_amd import _lib
from pyoomph_amd import jit_table as jt
from tests.oracle_bridge import oracle_problem, product_problem

pytestmark = pytest.mark.gpu

TH_SLOTS = [(l, v) for l in range(9) for v in ((0, 1, 2) if l in (0, 2, 6, 8) else (0, 1))]


def element_info(op, e, x, U_hist=None, aux=None):
    """JITElementInfo_t of element e of an oracle problem (values at time 0, and the
    history levels 1 and 2 when given)."""
    E = op.mesh.elements[e]
    nn = len(E)
    nf = op.nvalue + (1 if aux is not None else 0)
    U = op.nodal_values(x)
    ntime = 3
    vals = np.zeros((nn, nf, ntime))
    leq = np.full((nn, nf), -1, dtype=np.int32)
    for l in range(nn):
        for v in range(op.nvalue):
            vals[l, v, 0] = U[E[l], v]
            if U_hist is not None:
                vals[l, v, 1] = U_hist[0][E[l], v]
                vals[l, v, 2] = U_hist[1][E[l], v]
        if aux is not None:
            vals[l, op.nvalue, 0] = aux[E[l]]
    for s, (node, val) in enumerate(op.slots):
        leq[node, val] = op.leqn[e, s]
    return jt.ElementInfo(op.mesh.coords[E], vals, leq)


def call_slot(slot, info, nloc, flag, bdf2=None):
    R = np.zeros(max(nloc, 1))
    J = np.zeros((max(nloc, 1), max(nloc, 1)))
    M = np.zeros_like(J)
    si, keep = jt.shape_info(max(nloc, 1), bdf2)
    slot(ctypes.byref(info.info), ctypes.c_void_p(ctypes.addressof(si)), R.ctypes.data_as(jt.c_dp),
         J.ctypes.data_as(jt.c_dp), M.ctypes.data_as(jt.c_dp), flag)
    return R[:nloc], J[:nloc, :nloc], M[:nloc, :nloc]


CASES = [("poisson", dict(N=4, seed=3), _lib.ELEM_TRI_P2, True),
         ("bratu", dict(N=5, space="C1", lam=1.7), _lib.ELEM_TRI_P1, True),
         ("bratu", dict(N=4, space="C2", lam=2.3), _lib.ELEM_TRI_P2, True),
         ("rd", dict(N=2), _lib.ELEM_HEX_Q2, False),
         ("ns", dict(N=3, Re=50.0), _lib.ELEM_QUAD_Q2, False)]


@pytest.mark.parametrize("kind,kw,elem,inexact", CASES)
def test_steady_slots_equal_oracle_elements(gpu, kind, kw, elem, inexact):
    mp = product_problem(kind, **kw)
    okw = dict(kw)
    if kind == "rd":
        okw["f"] = mp.equations.source_field
    op = oracle_problem(kind, **okw)
    form = mp.equations.form
    t = jt.JitTable(form, elem)
    try:
        if kind == "bratu":
            t.set_global_parameter(0, kw["lam"])
        elif kind in ("poisson", "ns"):
            t.set_form_params(mp.equations.params())
        x = np.random.default_rng(5).uniform(-0.4, 0.4, op.ndof)
        Ro, Jo = op.element_outputs(x, 1)
        aux = mp.equations.source_field if kind == "rd" else None
        nel = op.mesh.elements.shape[0]
        for e in sorted({0, 1, nel // 2, nel - 1}):
            info = element_info(op, e, x, aux=aux)
            nv = int(op.nvar[e])
            for name in ("ResidualAndJacobianSteady", "ResidualAndJacobian"):
                R, J, _ = call_slot(t.slot(name), info, nv, 1)
                ro, jo = Ro[e, :nv], Jo[e].reshape(-1)[:nv * nv].reshape(nv, nv)
                if inexact:
                    assert np.max(np.abs(R - ro)) <= 1e-12 * max(1.0, np.max(np.abs(ro)))
                    assert np.max(np.abs(J - jo)) <= 1e-12 * max(1.0, np.max(np.abs(jo)))
                else:
                    assert np.array_equal(R, ro) and np.array_equal(J, jo), (e, name)
                R0, J0, _ = call_slot(t.slot(name), info, nv, 0)
                assert np.array_equal(R0, R) and not J0.any()
            if kind == "bratu":
                Rd, Jd = op.element_outputs(x, 1, deriv=1)
                R, J, _ = call_slot(t.parameter_derivative(), info, nv, 1)
                assert np.max(np.abs(R - Rd[e, :nv])) <= 1e-12 * max(1.0, np.max(np.abs(Rd[e])))
                jd = Jd[e].reshape(-1)[:nv * nv].reshape(nv, nv)
                assert np.max(np.abs(J - jd)) <= 1e-12 * max(1.0, np.max(np.abs(jd)))
    finally:
        t.close()


def test_navier_stokes_time_dependent_slot(gpu):
    """ResidualAndJacobian of Taylor-Hood with the shape info's BDF2 weights and the nodal
    history: equals the oracle's transient element (residual, Jacobian, mass matrix)."""
    mp = product_problem("ns", N=3, Re=40.0)
    op = oracle_problem("ns", N=3, Re=40.0)
    rng = np.random.default_rng(11)
    x, h1, h2 = (rng.uniform(-0.3, 0.3, op.ndof) for _ in range(3))
    tw = np.array([2.5, -3.1, 0.6])
    op.set_time(tw, h1, h2)
    Ro, Jo = op.element_outputs(x, 1)
    _, Mo = op.element_outputs(x, 2)
    t = jt.JitTable(mp.equations.form, _lib.ELEM_QUAD_Q2)
    try:
        t.set_form_params(mp.equations.params())
        hist = (op.nodal_values(h1), op.nodal_values(h2))
        for e in (0, 4, 8):
            info = element_info(op, e, x, U_hist=hist)
            nv = int(op.nvar[e])
            R, J, M = call_slot(t.slot("ResidualAndJacobian"), info, nv, 2, bdf2=tw)
            assert np.array_equal(R, Ro[e, :nv])
            assert np.array_equal(J, Jo[e].reshape(-1)[:nv * nv].reshape(nv, nv))
            assert np.array_equal(M, Mo[e].reshape(-1)[:nv * nv].reshape(nv, nv))
    finally:
        t.close()


def test_superlu_entry_solves_like_splu(gpu):
    op = oracle_problem("poisson", N=10, seed=1)
    x = np.zeros(op.ndof)
    r, rs, col, val = op.get_jacobian(x)
    n, nnz = op.ndof, len(val)
    L = _lib.load()
    ci = ctypes.c_int
    values = np.ascontiguousarray(val, dtype=np.float64)
    rowind = np.ascontiguousarray(col, dtype=np.int32)     # J's columns = J^T's rows
    colptr = np.ascontiguousarray(rs, dtype=np.int32)
    op_, n_, nnz_, tr, doc, info = ci(1), ci(n), ci(nnz), ci(1), ci(0), ci(-1)
    fac = ctypes.c_void_p()
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))  # noqa: E731
    rc = L.pyhip_superlu(ctypes.byref(op_), ctypes.byref(n_), ctypes.byref(nnz_), None,
                         values.ctypes.data_as(jt.c_dp), ip(rowind), ip(colptr), None, ctypes.byref(n_),
                         ctypes.byref(tr), ctypes.byref(doc), ctypes.byref(fac), ctypes.byref(info))
    assert rc == 0 and info.value == 0 and fac.value
    J = sp.csr_matrix((val, col, rs), shape=(n, n))
    for seed in (0, 1):
        b = np.random.default_rng(seed).normal(size=n)
        xb = b.copy()
        op_.value, nrhs = 2, ci(1)
        rc = L.pyhip_superlu(ctypes.byref(op_), ctypes.byref(n_), None, ctypes.byref(nrhs), None, None, None,
                             xb.ctypes.data_as(jt.c_dp), ctypes.byref(n_), ctypes.byref(tr), ctypes.byref(doc),
                             ctypes.byref(fac), ctypes.byref(info))
        assert rc == 0 and info.value == 0
        xs = sla.splu(J.tocsc()).solve(b)
        assert np.max(np.abs(xb - xs)) <= 1e-10 * np.max(np.abs(xs))
    op_.value = 3
    assert L.pyhip_superlu(ctypes.byref(op_), ctypes.byref(n_), None, None, None, None, None, None, None, None, None,
                           ctypes.byref(fac), ctypes.byref(info)) == 0
    assert not fac.value


GEN_CASES = [("bratu", dict(N=4, space="C2", lam=2.3), _lib.ELEM_TRI_P2),
             ("bratu", dict(N=5, space="C1", lam=1.7), _lib.ELEM_TRI_P1),
             ("rd", dict(N=2), _lib.ELEM_HEX_Q2)]


@pytest.mark.parametrize("kind,kw,elem", GEN_CASES)
def test_generated_form_tables_equal_the_compiled_ones(gpu, kind, kw, elem):
    """A run-time generated form's JIT table: fields, global parameter and slots; its slot
    outputs equal the compiled form's slot outputs bit for bit (R, J, dR/dlambda)."""
    from pyoomph_amd.codegen import library
    from pyoomph_amd.codegen.runtime import compile_form
    mp = product_problem(kind, **kw)
    okw = dict(kw)
    aux = None
    if kind == "rd":
        okw["f"] = aux = mp.equations.source_field
    op = oracle_problem(kind, **okw)
    gform = library.bratu() if kind == "bratu" else library.reaction_diffusion()
    gid = compile_form(gform, {_lib.ELEM_TRI_P1: "tri_p1", _lib.ELEM_TRI_P2: "tri_p2",
                               _lib.ELEM_HEX_Q2: "hex_q2"}[elem])
    tc, tg = jt.JitTable(mp.equations.form, elem), jt.JitTable(gid, elem)
    try:
        space = "numfields_C1" if elem == _lib.ELEM_TRI_P1 else "numfields_C2"
        names = "fieldnames_C1" if elem == _lib.ELEM_TRI_P1 else "fieldnames_C2"
        nf = tc.uint(space)
        assert tg.uint(space) == nf and tg.names(names, nf) == tc.names(names, nf)
        assert tg.uint("numglobal_params") == tc.uint("numglobal_params")
        if kind == "bratu":
            tc.set_global_parameter(0, kw["lam"])
            tg.set_global_parameter(0, kw["lam"])
        x = np.random.default_rng(8).uniform(-0.4, 0.4, op.ndof)
        nel = op.mesh.elements.shape[0]
        for e in sorted({0, nel // 2, nel - 1}):
            info = element_info(op, e, x, aux=aux)
            nv = int(op.nvar[e])
            for name in ("ResidualAndJacobianSteady", "ResidualAndJacobian"):
                a = call_slot(tc.slot(name), info, nv, 1)
                b = call_slot(tg.slot(name), info, nv, 1)
                assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (e, name)
            if kind == "bratu":
                a = call_slot(tc.parameter_derivative(), info, nv, 1)
                b = call_slot(tg.parameter_derivative(), info, nv, 1)
                assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), e
    finally:
        tc.close()
        tg.close()


def test_generated_form_without_a_compiled_counterpart(gpu):
    """Nonlinear diffusion -div((1 + k u^2) grad u) = s0 through its JIT table: the slot's
    Jacobian is the central difference of the slot's residual (and not symmetric)."""
    from pyoomph_amd.codegen import library
    from pyoomph_amd.codegen.runtime import compile_form
    form = library.nonlinear_diffusion()
    gid = compile_form(form, "tri_p2")
    op = oracle_problem("bratu", N=3, space="C2", lam=1.0)     # mesh, numbering, slots
    t = jt.JitTable(gid, _lib.ELEM_TRI_P2)
    try:
        t.set_form_params([2.0, 8.0])
        rng = np.random.default_rng(2)
        x = rng.uniform(-0.5, 0.5, op.ndof)
        e = 4
        nv = int(op.nvar[e])
        R, J, _ = call_slot(t.slot("ResidualAndJacobianSteady"), element_info(op, e, x), nv, 1)
        d = rng.normal(size=op.ndof)
        eps = 1e-6
        Rp, _, _ = call_slot(t.slot("ResidualAndJacobianSteady"), element_info(op, e, x + eps * d), nv, 0)
        Rm, _, _ = call_slot(t.slot("ResidualAndJacobianSteady"), element_info(op, e, x - eps * d), nv, 0)
        leq = op.leqn[e]
        dl = np.zeros(nv)
        for s in range(len(leq)):
            if leq[s] >= 0:
                dl[leq[s]] = d[op.slot_geqn[e, s]]
        fd = (Rp - Rm) / (2 * eps)
        assert np.max(np.abs(J @ dl - fd)) <= 1e-6 * max(np.max(np.abs(fd)), 1e-12)
        assert not np.allclose(J, J.T)
    finally:
        t.close()
