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
