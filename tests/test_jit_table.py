"""The JIT element-table boundary on the CPU.

* Layout: every struct of include/pyoomph_jit_table.h has the size, and every field the
  offset, of the reference's src/jitbridge.h (a probe program compiled against each; the
  reference half runs where the reference tree is present).
* pyhip_jit_element_init fills the table as a generated JIT_ELEMENT_init does
  (src/codegen.cpp:6398-7090): field counts and names per space, one residual with its
  three slots, Bratu's lambda with its parameter-derivative slot, the dominant space, the
  default integration order, clean_up -- which frees and clears what the init allocated.
* The superlu() entry rejects bad calls through *info (oomph throws on a non-zero info,
  linear_solver.cc:2312-2320) and frees a null factorisation.
"""
import ctypes
import os
import sys

import pytest

from pyoomph_amd import _lib
from pyoomph_amd import jit_table as jt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import jit_layout  # noqa: E402


def test_own_layout_matches_the_python_mirror():
    own = jit_layout.layout("own")
    assert len(own) > 300
    for k, v in jt.FIELD_OFFSETS.items():
        assert own[k] == v, k
    assert own["JITShapeInfo_t sizeof"] == jt.SHAPEINFO_SIZE
    assert own["JITFuncSpec_Table_FiniteElement_t sizeof"] == jt.TABLE_SIZE
    assert own["JITElementInfo_t sizeof"] == ctypes.sizeof(jt.JITElementInfo)
    for f, _ in jt.JITElementInfo._fields_:
        assert own["JITElementInfo_t." + f] == getattr(jt.JITElementInfo, f).offset, f


@pytest.mark.skipif(not os.path.exists(jit_layout.REF_HEADER), reason="reference tree not present")
def test_layout_equals_the_reference_jitbridge():
    own, ref = jit_layout.layout("own"), jit_layout.layout("ref")
    assert set(own) == set(ref)
    assert {k: v for k, v in own.items() if ref[k] != v} == {}


CASES = [(_lib.FORM_POISSON, _lib.ELEM_TRI_P2, 2, {"C2": ["u"]}, "C2"),
         (_lib.FORM_BRATU, _lib.ELEM_TRI_P1, 2, {"C1": ["u"]}, "C1"),
         (_lib.FORM_BRATU, _lib.ELEM_TRI_P2, 2, {"C2": ["u"]}, "C2"),
         (_lib.FORM_REACTION_DIFFUSION, _lib.ELEM_HEX_Q2, 3, {"C2": ["u", "f"]}, "C2"),
         (_lib.FORM_NAVIER_STOKES_TH, _lib.ELEM_QUAD_Q2, 2,
          {"C2": ["velocity_x", "velocity_y"], "C1": ["pressure"]}, "C2")]


@pytest.mark.parametrize("form,kind,dim,fields,dominant", CASES)
def test_init_fills_the_table(form, kind, dim, fields, dominant):
    t = jt.JitTable(form, kind)
    assert t.uint("nodal_dim") == dim
    for sp in ("C1", "C2"):
        n = t.uint("numfields_" + sp)
        assert n == len(fields.get(sp, []))
        assert t.names("fieldnames_" + sp, n) == fields.get(sp, [])
    assert t.uint("num_res_jacs") == 1 and t.names("res_jac_names", 1) == [""]
    for s in ("ResidualAndJacobian", "ResidualAndJacobianSteady", "ResidualAndJacobian_NoHang"):
        assert t.slot(s) is not None
    if form == _lib.FORM_BRATU:
        assert t.uint("numglobal_params") == 1 and t.parameter_derivative() is not None
    else:
        assert t.uint("numglobal_params") == 0 and t.pointer("ParameterDerivative") is None
    assert t.string("dominant_space") == dominant
    assert t.int("integration_order") == 0
    t.close()
    for f in ("ResidualAndJacobian", "ResidualAndJacobianSteady", "res_jac_names", "fieldnames_C2",
              "ParameterDerivative", "dominant_space"):
        assert t.pointer(f) is None, f


def test_init_rejects_unknown_combinations():
    L = _lib.load()
    jt._signatures(L)
    buf = ctypes.create_string_buffer(jt.TABLE_SIZE)
    rc = L.pyhip_jit_element_init(_lib.FORM_NAVIER_STOKES_TH, _lib.ELEM_TRI_P2, ctypes.c_void_p(ctypes.addressof(buf)))
    assert rc == _lib.PYHIP_ERR_INVALID
    assert L.pyhip_jit_element_init(9, _lib.ELEM_TRI_P2, ctypes.c_void_p(ctypes.addressof(buf))) == _lib.PYHIP_ERR_INVALID
    t = jt.JitTable(_lib.FORM_POISSON, _lib.ELEM_TRI_P1)
    t.close()
    with pytest.raises(RuntimeError, match="not filled"):
        t.set_form_params([1.0])


def test_superlu_entry_error_paths():
    L = _lib.load()
    L.pyhip_superlu.restype = ctypes.c_int
    handle = ctypes.c_void_p()
    info = ctypes.c_int(-7)
    op = ctypes.c_int(3)
    n = ctypes.c_int(4)
    # freeing an empty factorisation is fine
    assert L.pyhip_superlu(ctypes.byref(op), ctypes.byref(n), None, None, None, None, None, None, None, None, None,
                           ctypes.byref(handle), ctypes.byref(info)) == 0
    assert info.value == 0
    # an unknown op is an error reported through info
    op.value = 7
    rc = L.pyhip_superlu(ctypes.byref(op), ctypes.byref(n), None, None, None, None, None, None, None, None, None,
                         ctypes.byref(handle), ctypes.byref(info))
    assert rc == _lib.PYHIP_ERR_INVALID and info.value == -_lib.PYHIP_ERR_INVALID
    # a solve before any factorisation too
    op.value = 2
    b = (ctypes.c_double * 4)()
    rc = L.pyhip_superlu(ctypes.byref(op), ctypes.byref(n), None, None, None, None, None, b, ctypes.byref(n), None,
                         None, ctypes.byref(handle), ctypes.byref(info))
    assert rc != 0 and info.value != 0
