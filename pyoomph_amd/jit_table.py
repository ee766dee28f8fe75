"""Host-side mirror of the JIT element-table ABI (include/pyoomph_jit_table.h).

``JitTable(form, kind)`` allocates a zeroed JITFuncSpec_Table_FiniteElement_t the way the
reference host does (src/problem.cpp:110-111), has ``pyhip_jit_element_init`` fill it (the
role of a generated JIT_ELEMENT_init, src/codegen.cpp:6398-7090), reads its fields and calls
its per-element slots with a ``JITElementInfo_t`` built from numpy arrays (the role of
BulkElementBase::fill_element_info, src/elements.cpp:2713-3020). ``close()`` calls the
table's clean_up, as the host does before deleting it (src/problem.cpp:144-165).

Field offsets come from the C layout: ctypes structures below for the element info and the
few shape-info fields the slots read, and the table read through ``FIELD_OFFSETS``, the
offsets of the fields this module reads, as the C compiler lays them out (the CPU tests
check both against a probe compiled from the header).
"""
import ctypes

import numpy as np

from . import _lib

c_dp = ctypes.POINTER(ctypes.c_double)
c_dpp = ctypes.POINTER(c_dp)
c_dppp = ctypes.POINTER(c_dpp)
c_ip = ctypes.POINTER(ctypes.c_int)
c_ipp = ctypes.POINTER(c_ip)


class JITElementInfo(ctypes.Structure):
    pass


JITElementInfo._fields_ = [
    ("nnode", ctypes.c_uint), ("nnode_C1", ctypes.c_uint), ("nnode_C2", ctypes.c_uint),
    ("nnode_C1TB", ctypes.c_uint), ("nnode_C2TB", ctypes.c_uint), ("nnode_DL", ctypes.c_uint),
    ("nodal_dim", ctypes.c_uint), ("nodal_coords", c_dppp), ("nodal_data", c_dppp),
    ("nodal_local_eqn", c_ipp), ("pos_local_eqn", c_ipp), ("ndof", ctypes.c_uint),
    ("alloced", ctypes.c_bool), ("elem_ptr", ctypes.c_void_p),
    ("bulk_eleminfo", ctypes.POINTER(JITElementInfo)), ("opposite_eleminfo", ctypes.POINTER(JITElementInfo))]

# sizes and offsets of the C layout (x86-64 / LP64) for what this module touches
SHAPEINFO_SIZE = 624
TABLE_SIZE = 1200
FIELD_OFFSETS = {
    "JITShapeInfo_t.jacobian_size": 448,
    "JITShapeInfo_t.timestepper_weights_dt_BDF2": 512,
    "JITFuncSpec_Table_FiniteElement_t.nodal_dim": 0,
    "JITFuncSpec_Table_FiniteElement_t.numfields_C1": 8,
    "JITFuncSpec_Table_FiniteElement_t.fieldnames_C1": 24,
    "JITFuncSpec_Table_FiniteElement_t.numfields_C2": 32,
    "JITFuncSpec_Table_FiniteElement_t.fieldnames_C2": 48,
    "JITFuncSpec_Table_FiniteElement_t.num_res_jacs": 444,
    "JITFuncSpec_Table_FiniteElement_t.res_jac_names": 456,
    "JITFuncSpec_Table_FiniteElement_t.numglobal_params": 528,
    "JITFuncSpec_Table_FiniteElement_t.global_parameters": 544,
    "JITFuncSpec_Table_FiniteElement_t.ParameterDerivative": 552,
    "JITFuncSpec_Table_FiniteElement_t.dominant_space": 624,
    "JITFuncSpec_Table_FiniteElement_t.integration_order": 652,
    "JITFuncSpec_Table_FiniteElement_t.ResidualAndJacobian": 672,
    "JITFuncSpec_Table_FiniteElement_t.ResidualAndJacobianSteady": 680,
    "JITFuncSpec_Table_FiniteElement_t.ResidualAndJacobian_NoHang": 688,
    "JITFuncSpec_Table_FiniteElement_t.hessian_generated": 712,
    "JITFuncSpec_Table_FiniteElement_t.clean_up": 1192,
}

SLOT = ctypes.CFUNCTYPE(None, ctypes.POINTER(JITElementInfo), ctypes.c_void_p, c_dp, c_dp, c_dp, ctypes.c_uint)
CLEAN_UP = ctypes.CFUNCTYPE(None, ctypes.c_void_p)


def _signatures(L):
    L.pyhip_jit_element_init.restype = ctypes.c_int
    L.pyhip_jit_element_init.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.pyhip_jit_set_form_params.restype = ctypes.c_int
    L.pyhip_jit_set_form_params.argtypes = [ctypes.c_void_p, c_dp, ctypes.c_int]


class JitTable:
    """One filled element table; fields read by name, slots called per element."""

    def __init__(self, form, kind):
        L = _lib.load()
        _signatures(L)
        self._buf = ctypes.create_string_buffer(TABLE_SIZE)     # zeroed, like the host's new
        self.addr = ctypes.addressof(self._buf)
        _lib.check(L.pyhip_jit_element_init(int(form), int(kind), ctypes.c_void_p(self.addr)))
        self.form, self.kind = form, kind
        self._param_store = None

    def _at(self, field, ctype):
        return ctype.from_address(self.addr + FIELD_OFFSETS["JITFuncSpec_Table_FiniteElement_t." + field])

    def uint(self, field):
        return self._at(field, ctypes.c_uint).value

    def int(self, field):
        return self._at(field, ctypes.c_int).value

    def pointer(self, field):
        return self._at(field, ctypes.c_void_p).value

    def names(self, field, count):
        p = self._at(field, ctypes.POINTER(ctypes.c_char_p))
        return [p[i].decode() for i in range(count)] if p else []

    def string(self, field):
        v = self._at(field, ctypes.c_char_p).value
        return None if v is None else v.decode()

    def slot(self, field, res=0):
        arr = self._at(field, ctypes.POINTER(ctypes.c_void_p))
        return SLOT(arr[res]) if arr and arr[res] else None

    def parameter_derivative(self, res=0, param=0):
        pd = self._at("ParameterDerivative", ctypes.POINTER(ctypes.POINTER(ctypes.c_void_p)))
        return SLOT(pd[res][param]) if pd and pd[res] and pd[res][param] else None

    def set_global_parameter(self, i, value):
        """Points global_parameters[i] at a value this object keeps (the host's storage)."""
        if self._param_store is None:
            self._param_store = (ctypes.c_double * 8)()
        self._param_store[i] = value
        gp = self._at("global_parameters", ctypes.POINTER(c_dp))
        gp[i] = ctypes.cast(ctypes.byref(self._param_store, 8 * i), c_dp)

    def set_form_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        _lib.check(_lib.load().pyhip_jit_set_form_params(ctypes.c_void_p(self.addr), p.ctypes.data_as(c_dp),
                                                         int(p.size)))

    def close(self):
        fn = self.pointer("clean_up")
        if fn:
            CLEAN_UP(fn)(ctypes.c_void_p(self.addr))


class ElementInfo:
    """JITElementInfo_t of one element from numpy arrays: coords [nnode, dim], values
    [nnode, nfield, ntime] and local equations [nnode, nfield] (-1 pinned)."""

    def __init__(self, coords, values, local_eqn):
        coords = np.asarray(coords, dtype=np.float64)
        values = np.asarray(values, dtype=np.float64)
        local_eqn = np.asarray(local_eqn, dtype=np.int32)
        nn, dim = coords.shape
        self._keep = []
        self._c = np.ascontiguousarray(coords[:, :, None])          # time index 0 only
        self._v = np.ascontiguousarray(values)
        self._e = np.ascontiguousarray(local_eqn)

        def ppp(arr):
            n0, n1 = arr.shape[0], arr.shape[1]
            rows = (c_dpp * n0)()
            for i in range(n0):
                cols = (c_dp * n1)()
                for j in range(n1):
                    cols[j] = arr[i, j].ctypes.data_as(c_dp)
                self._keep.append(cols)
                rows[i] = cols
            self._keep.append(rows)
            return rows

        eq_rows = (c_ip * nn)()
        for i in range(nn):
            eq_rows[i] = self._e[i].ctypes.data_as(c_ip)
        self._keep.append(eq_rows)
        self.info = JITElementInfo()
        self.info.nnode = nn
        self.info.nodal_dim = dim
        self.info.nodal_coords = ppp(self._c)
        self.info.nodal_data = ppp(self._v)
        self.info.nodal_local_eqn = eq_rows
        self.info.ndof = int(max(local_eqn.max() + 1, 0))


def shape_info(jacobian_size, bdf2_weights=None):
    """A zeroed JITShapeInfo_t with the fields the slots read."""
    buf = ctypes.create_string_buffer(SHAPEINFO_SIZE)
    ctypes.c_uint.from_address(ctypes.addressof(buf) + FIELD_OFFSETS["JITShapeInfo_t.jacobian_size"]).value = \
        int(jacobian_size)
    keep = [buf]
    if bdf2_weights is not None:
        w = (ctypes.c_double * 3)(*[float(x) for x in bdf2_weights])
        keep.append(w)
        ctypes.c_void_p.from_address(ctypes.addressof(buf) +
                                     FIELD_OFFSETS["JITShapeInfo_t.timestepper_weights_dt_BDF2"]).value = \
            ctypes.addressof(w)
    return buf, keep
