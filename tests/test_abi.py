"""C ABI: the library loads without a GPU and exports every entry point the header
declares; host-only entry points behave."""
import ctypes
import os
import re

import pytest

from pyoomph_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    inc = os.path.join(ROOT, "include")
    text = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc)) if f.endswith(".h"))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pyhip_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = _lib.load()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, "binding missing for " + n


def test_element_table():
    L = _lib.load()
    t = _lib.ElementTable()
    assert L.pyhip_element_init(_lib.FORM_NAVIER_STOKES_TH, _lib.ELEM_QUAD_Q2, ctypes.byref(t)) == 0
    assert (t.nslot, t.nnode, t.nqp, t.nodal_values, t.dim) == (22, 9, 9, 3, 2)
    assert L.pyhip_element_init(_lib.FORM_BRATU, _lib.ELEM_TRI_P2, ctypes.byref(t)) == 0
    assert (t.nslot, t.nqp, t.numglobal_params, t.has_param_derivative[0]) == (6, 7, 1, 1)
    assert t.global_param_names[0] == b"lambda"
    assert L.pyhip_element_init(_lib.FORM_POISSON, _lib.ELEM_HEX_Q2, ctypes.byref(t)) == 0
    assert (t.nslot, t.nqp) == (27, 27)
    assert L.pyhip_element_init(_lib.FORM_POISSON, _lib.ELEM_TRI_P1, ctypes.byref(t)) == 0
    assert t.nqp == 3


def test_invalid_arguments_report_errors():
    L = _lib.load()
    t = _lib.ElementTable()
    rc = L.pyhip_element_init(_lib.FORM_NAVIER_STOKES_TH, _lib.ELEM_TRI_P2, ctypes.byref(t))
    assert rc == _lib.PYHIP_ERR_INVALID
    assert b"Taylor-Hood" in L.pyhip_last_error()
    with pytest.raises(RuntimeError):
        _lib.check(L.pyhip_element_init(99, 1, ctypes.byref(t)))


def test_default_options():
    o = _lib.NewtonOpts()
    _lib.load().pyhip_newton_opts_default(ctypes.byref(o))
    assert o.tol == 1e-8 and o.max_iter == 10 and o.max_residuals == 1e10
    assert o.always_take_one_newton_step == 1 and o.relaxation == 1.0


def _plan(reqs, nhvec=0):
    """pyhip_multi_assembly_plan over (what, param, hvec, transposed) tuples."""
    L = _lib.load()
    arr = (_lib.MultiRequest * max(len(reqs), 1))()
    for i, (w, p, h, t) in enumerate(reqs):
        arr[i].what, arr[i].param, arr[i].hvec, arr[i].transposed = w, p, h, t
    ret = (ctypes.c_int * max(len(reqs), 1))()
    nv, nm = ctypes.c_int(), ctypes.c_int()
    rc = L.pyhip_multi_assembly_plan(len(reqs), arr, nhvec, ret, ctypes.byref(nv), ctypes.byref(nm))
    if rc:
        return rc, L.pyhip_last_error().decode()
    return [ret[i] for i in range(len(reqs))], (nv.value, nm.value)


def test_multi_assembly_numbering():
    """Vectors count 0, 1, ..., matrices -1, -2, ... in request order, as
    CustomMultiAssembleHandler numbers one contribution (src/bifurcation.cpp:5393-5480),
    e.g. the Hopf tracker's R, dR/dp, J, J, M request (bifurcation_tools.py:1103)."""
    W = _lib.MULTI_WHAT
    reqs = [(W["residuals"][0], -1, -1, 0), (W["dresiduals_dparameter"][0], 0, -1, 0),
            (W["jacobian"][0], -1, -1, 0), (W["mass_matrix"][0], -1, -1, 0),
            (W["hessian_vector_product"][0], -1, 0, 0), (W["djacobian_dparameter"][0], 0, -1, 0),
            (W["mass_matrix_hessian_vector_product"][0], -1, 0, 0),
            (W["hessian_vector_product"][0], -1, 1, 0)]
    ret, counts = _plan(reqs, nhvec=2)
    assert ret == [0, 1, -1, -2, -3, -4, -5, -6] and counts == (2, 6)
    assert _plan([]) == ([], (0, 0))


@pytest.mark.parametrize("reqs,nh,msg", [
    ([(0, -1, -1, 0), (0, -1, -1, 0)], 0, "multiple 'residuals'"),
    ([(1, -1, -1, 0), (2, -1, -1, 0), (1, -1, -1, 0)], 0, "multiple 'jacobian'"),
    ([(4, 0, -1, 0), (4, 0, -1, 0)], 0, "multiple 'djacobian_dparameter'"),
    ([(6, -1, 0, 0), (6, -1, 0, 0)], 1, "multiple 'hessian_vector_product'"),
    ([(6, -1, 0, 0), (6, -1, 1, 1)], 2, "transposed and non-transposed"),
    ([(6, -1, 2, 0)], 2, "out of bounds"),
    ([(3, -1, -1, 0)], 0, "no parameter"),
    ([(7, -1, -1, 0)], 1, "no Hessian vector"),
    ([(9, -1, -1, 0)], 0, "unknown multi-assembly request"),
])
def test_multi_assembly_request_errors(reqs, nh, msg):
    rc, err = _plan(reqs, nh)
    assert rc == _lib.PYHIP_ERR_INVALID and msg in err


def test_multi_assembly_distinct_requests_pass():
    """The same kind with another parameter, vector or transposition is a new quantity."""
    ret, counts = _plan([(6, -1, 0, 0), (6, -1, 1, 0), (7, -1, 0, 0)], 2)
    assert ret == [-1, -2, -3] and counts == (0, 3)
    ret, counts = _plan([(6, -1, 0, 1), (7, -1, 0, 1)], 1)
    assert ret == [-1, -2]


def test_multi_assembly_numbering_per_contribution():
    """Contributions (residual tables) are numbered one after the other in order of first
    appearance, each in request order (src/bifurcation.cpp:5436-5480)."""
    L = _lib.load()
    reqs = [(0, ""), (1, "mass"), (1, ""), (0, "mass")]
    arr = (_lib.MultiRequest * 4)()
    keep = []
    for i, (w, c) in enumerate(reqs):
        arr[i].what, arr[i].param, arr[i].hvec, arr[i].transposed = w, -1, -1, 0
        keep.append(c.encode())
        arr[i].contribution = keep[-1]
    ret = (ctypes.c_int * 4)()
    nv, nm = ctypes.c_int(), ctypes.c_int()
    assert L.pyhip_multi_assembly_plan(4, arr, 0, ret, ctypes.byref(nv), ctypes.byref(nm)) == 0
    assert list(ret) == [0, -2, -1, 1] and (nv.value, nm.value) == (2, 2)
    # the same quantity in two contributions is two quantities; twice in one is an error
    arr[3].what, arr[3].contribution = 1, keep[0]
    assert L.pyhip_multi_assembly_plan(4, arr, 0, ret, ctypes.byref(nv), ctypes.byref(nm)) == _lib.PYHIP_ERR_INVALID
    assert b"for the same contribution ''" in L.pyhip_last_error()
