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
