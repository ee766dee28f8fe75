"""Compiling generated element kernels into the running library, and using them.

``GeneratedEquation(form, space=..., params=...)`` is an equation object like the built-in
ones (pyoomph_amd.equations): a ``MeshProblem`` on a mesh of the form's element family
with it assembles and solves through the generated kernels. The first use of a (form,
element family) pair emits the kernel source (``scalar.emit_kernels``), compiles it with
hiprtc inside libpyoomph_hip.so and loads it on the GPU (``pyhip_codegen_compile``); later
uses take the registered id -- the reference likewise compiles each element's code once
and keeps the loaded library (src/ccompiler.cpp:191-248).
"""
import ctypes
import os

import numpy as np

from .. import _lib
from ..equations import Equations
from . import face, mixed, scalar

_HERE = os.path.dirname(os.path.abspath(__file__))
COMMON_HEADER = os.path.join(os.path.dirname(_HERE), "csrc", "element_common.h")

_registry = {}   # (form name, kind id) -> form id


def _signatures(L):
    L.pyhip_codegen_compile.restype = ctypes.c_int
    L.pyhip_codegen_compile.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p),
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_size_t)]


def compile_form(form, kind, load=True):
    """Compiles the form's kernels for element family ``kind`` with hiprtc; with load, also
    loads them on the current device and returns the form id (cached per form name and
    family). Without load (no GPU needed) returns the code-object size."""
    kid = scalar.KIND_IDS[kind] if isinstance(kind, str) else int(kind)
    key = (form.name, kid)
    if load and key in _registry:
        return _registry[key]
    th = isinstance(form, mixed.TaylorHoodWeakForm)
    if th:
        if kid != scalar.KIND_IDS["quad_q2"]:
            raise ValueError("Taylor-Hood forms are generated for Q2 quadrilaterals")
        src, name, info = mixed.emit_th_kernels(form)
    else:
        src, name, info = scalar.emit_kernels(form, kid)
    L = _lib.load()
    _signatures(L)
    with open(COMMON_HEADER) as f:
        hdr = f.read()
    hs = (ctypes.c_char_p * 1)(hdr.encode())
    hn = (ctypes.c_char_p * 1)(b"element_common.h")
    fid = ctypes.c_int(-1)
    size = ctypes.c_size_t(0)
    _lib.check(L.pyhip_codegen_compile(name.encode(), src.encode(), 1, hs, hn, kid, int(info["symmetric"]),
                                       int(info["has_param_derivative"]), int(bool(load)), ctypes.byref(fid),
                                       ctypes.byref(size)))
    if not load:
        return size.value
    if th:
        # velocity_x, velocity_y at every node, pressure at the corners
        _lib.check(L.pyhip_codegen_set_layout(fid.value, 1))
    # the names its JIT element table lists: the unknown "u", the auxiliary field "f"
    # (Taylor-Hood tables list velocity and pressure)
    _lib.check(L.pyhip_codegen_set_fields(fid.value, b"u", b"f" if form.aux else None,
                                          int(form.global_param or 0)))
    # named residual tables: each a generated form of its own, attached under its name
    for k, (tname, (tflux, tsource)) in enumerate(getattr(form, "tables", {}).items()):
        tform = scalar.ScalarWeakForm("%s__table%d" % (form.name, k), tflux, tsource, nparams=form.nparams,
                                      global_param=form.global_param,
                                      global_param_name=form.global_param_name or "lambda", aux=form.aux)
        tid = compile_form(tform, kid)
        _lib.check(L.pyhip_codegen_add_residual_table(fid.value, tname.encode(), tid))
    _registry[key] = fid.value
    return fid.value


def compile_constraint_form(form, kind, load=True):
    """Compiles an EnforcedBC constraint form (face.ConstraintForm) for the line faces of
    element family ``kind`` and, with load, registers it (pyhip_codegen_set_constraint_form);
    returns its id (cached per form name and family), or the code-object size without load."""
    kid = scalar.KIND_IDS[kind] if isinstance(kind, str) else int(kind)
    key = ("constraint", form.name, kid)
    if load and key in _registry:
        return _registry[key]
    src, name, info = face.emit_constraint_kernels(form, kid)
    L = _lib.load()
    _signatures(L)
    with open(COMMON_HEADER) as f:
        hdr = f.read()
    hs = (ctypes.c_char_p * 1)(hdr.encode())
    hn = (ctypes.c_char_p * 1)(b"element_common.h")
    fid = ctypes.c_int(-1)
    size = ctypes.c_size_t(0)
    _lib.check(L.pyhip_codegen_compile(name.encode(), src.encode(), 1, hs, hn, kid, 0,
                                       int(info["has_param_derivative"]), int(bool(load)), ctypes.byref(fid),
                                       ctypes.byref(size)))
    if not load:
        return size.value
    _lib.check(L.pyhip_codegen_set_constraint_form(fid.value, form.ncoef + (1 if info["normal"] else 0)))
    _lib.check(L.pyhip_codegen_set_fields(fid.value, b"u", None, int(form.global_param or 0)))
    _registry[key] = fid.value
    return fid.value


def compile_face_form(form, kind, load=True):
    """Compiles a boundary-face form (face.FaceFluxForm) for the line faces of element family
    ``kind`` and, with load, registers it as a face form (pyhip_codegen_set_face_form);
    returns its id (cached per form name and family), or the code-object size without load."""
    kid = scalar.KIND_IDS[kind] if isinstance(kind, str) else int(kind)
    key = ("face", form.name, kid)
    if load and key in _registry:
        return _registry[key]
    src, name, info = face.emit_face_kernels(form, kid)
    L = _lib.load()
    _signatures(L)
    with open(COMMON_HEADER) as f:
        hdr = f.read()
    hs = (ctypes.c_char_p * 1)(hdr.encode())
    hn = (ctypes.c_char_p * 1)(b"element_common.h")
    fid = ctypes.c_int(-1)
    size = ctypes.c_size_t(0)
    _lib.check(L.pyhip_codegen_compile(name.encode(), src.encode(), 1, hs, hn, kid, 0,
                                       int(info["has_param_derivative"]), int(bool(load)), ctypes.byref(fid),
                                       ctypes.byref(size)))
    if not load:
        return size.value
    _lib.check(L.pyhip_codegen_set_face_form(fid.value, form.ncoef + (1 if info["normal"] else 0)))
    _lib.check(L.pyhip_codegen_set_fields(fid.value, b"u", None, int(form.global_param or 0)))
    _registry[key] = fid.value
    return fid.value


class GeneratedEquation(Equations):
    """A generated weak form on one element family: scalar forms (ScalarWeakForm) on any
    family, Taylor-Hood forms (mixed.TaylorHoodWeakForm) on Q2 quadrilaterals.

    space: "C1" (P1 triangles) or "C2" (P2 triangles, Q2 quadrilaterals, Q2 bricks); the
    element family is the mesh's, known when the problem is built (``check_mesh``).
    params: the form's parameter values (``form.nparams`` of them); aux: the nodal field f
    (forms with aux=True)."""

    def __init__(self, form, *, space="C2", params=(), aux=None):
        self.gen_form = form
        self.space = space
        self.taylor_hood = isinstance(form, mixed.TaylorHoodWeakForm)
        if self.taylor_hood:
            self.nvalue = 3
        self._params = np.asarray(params, dtype=np.float64).reshape(-1)
        if self._params.size != form.nparams:
            raise ValueError("the form takes %d parameters" % form.nparams)
        self._aux = None if aux is None else np.asarray(aux, dtype=np.float64)
        if form.aux and self._aux is None:
            raise ValueError("the form reads the auxiliary field: pass aux")
        self.global_parameters = (form.global_param_name,) if form.global_param is not None else ()
        self._kind = None
        self.symmetric = False

    def check_mesh(self, mesh):
        super().check_mesh(mesh)
        if self.taylor_hood and mesh.kind != "quad_q2":
            raise ValueError("Taylor-Hood needs a Q2 quadrilateral mesh")
        if self._kind is not None and self._kind != scalar.KIND_IDS[mesh.kind]:
            raise ValueError("one GeneratedEquation object per element family")
        self._kind = scalar.KIND_IDS[mesh.kind]
        self.symmetric = self.gen_form.symmetric(scalar.KIND_DIM[self._kind])

    def dummy_values(self, mesh):
        """Taylor-Hood: the pressure lives on the element vertices; the other nodes carry a
        pinned dummy (as NavierStokesEquations)."""
        if not self.taylor_hood:
            return None
        vertex = np.zeros(mesh.nnode, dtype=bool)
        vertex[mesh.elements[:, [0, 2, 6, 8]].reshape(-1)] = True
        d = np.zeros((mesh.nnode, 3), dtype=bool)
        d[~vertex, 2] = True
        return d

    @property
    def form(self):
        """The registered form id (compiles and loads the kernels on first use, on the
        current device)."""
        if self._kind is None:
            raise RuntimeError("the equation has not been put on a mesh yet")
        return compile_form(self.gen_form, self._kind)

    def params(self):
        return self._params.copy()

    def aux(self, mesh):
        if self._aux is None:
            return None
        if self._aux.shape != (mesh.nnode,):
            raise ValueError("aux must have one value per node")
        return self._aux
