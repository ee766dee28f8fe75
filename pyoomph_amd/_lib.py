"""ctypes binding of libpyoomph_hip.so (the C ABI declared in include/pyoomph_hip.h).

The library is built in-tree (``pyoomph_amd/libpyoomph_hip.so``). Loading fails loudly
when it is missing: there is no CPU fallback. ``PYOOMPH_HIP_LIB`` names another build of
the same sources instead (bench.py's assembly study loads one compiled with FP
contraction this way, in a process of its own).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PYOOMPH_HIP_LIB") or os.path.join(_HERE, "libpyoomph_hip.so")

PYHIP_OK = 0
PYHIP_ERR_INVALID = -1
PYHIP_ERR_HIP = -2
PYHIP_ERR_NOMEM = -3
PYHIP_ERR_NOT_CONVERGED = -4
PYHIP_ERR_BREAKDOWN = -5
PYHIP_ERR_NEWTON = -6

ELEM_TRI_P1, ELEM_TRI_P2, ELEM_QUAD_Q2, ELEM_HEX_Q2 = 1, 2, 3, 4
ELEM_KIND = {"tri_p1": ELEM_TRI_P1, "tri_p2": ELEM_TRI_P2, "quad_q2": ELEM_QUAD_Q2,
             "hex_q2": ELEM_HEX_Q2}
FORM_POISSON, FORM_BRATU, FORM_REACTION_DIFFUSION, FORM_NAVIER_STOKES_TH = 1, 2, 3, 4
F64, F32 = 0, 1
PRECOND_JACOBI, PRECOND_ILU0, PRECOND_AMG, PRECOND_BLOCK = 0, 1, 2, 3
ILU_ORDER_AUTO, ILU_ORDER_NATURAL, ILU_ORDER_MULTICOLOR = 0, 1, 2
KRYLOV_GMRES, KRYLOV_CG = 0, 1

c_i32p = ctypes.POINTER(ctypes.c_int32)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_dp = ctypes.POINTER(ctypes.c_double)


class SubproblemDesc(ctypes.Structure):
    _fields_ = [("form", ctypes.c_int), ("elem_kind", ctypes.c_int),
                ("nnode", ctypes.c_int64), ("nelem", ctypes.c_int64),
                ("coords", c_dp), ("conn", c_i32p), ("nvalue", ctypes.c_int),
                ("eqn", c_i32p), ("pinned_values", c_dp), ("ndof", ctypes.c_int64),
                ("params", c_dp), ("nparams", ctypes.c_int), ("aux", c_dp),
                ("nface", ctypes.c_int64), ("face_conn", c_i32p), ("face_coef", c_dp),
                ("nenforced", ctypes.c_int64), ("enforced_conn", c_i32p), ("enforced_coef", c_dp),
                ("face_form", ctypes.c_int), ("enforced_form", ctypes.c_int)]


class ElementTable(ctypes.Structure):
    _fields_ = [("form", ctypes.c_int), ("elem_kind", ctypes.c_int),
                ("num_res_jacs", ctypes.c_int), ("res_jac_names", ctypes.c_char_p * 2),
                ("nodal_values", ctypes.c_int), ("nslot", ctypes.c_int),
                ("nnode", ctypes.c_int), ("dim", ctypes.c_int),
                ("integration_order", ctypes.c_int), ("nqp", ctypes.c_int),
                ("numglobal_params", ctypes.c_int),
                ("global_param_names", ctypes.c_char_p * 2),
                ("has_param_derivative", ctypes.c_int * 2), ("moving_nodes", ctypes.c_int)]


class LinearOpts(ctypes.Structure):
    _fields_ = [("krylov", ctypes.c_int), ("precond", ctypes.c_int), ("precision", ctypes.c_int),
                ("restart", ctypes.c_int), ("max_iters", ctypes.c_int),
                ("rtol", ctypes.c_double), ("rtol_f32", ctypes.c_double),
                ("max_refinements", ctypes.c_int),
                ("refine_rtol", ctypes.c_double), ("ilu_order", ctypes.c_int), ("amg_element_level", ctypes.c_int)]


class NewtonOpts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("max_residuals", ctypes.c_double),
                ("always_take_one_newton_step", ctypes.c_int),
                ("relaxation", ctypes.c_double), ("jacobian_precision", ctypes.c_int),
                ("linear", LinearOpts), ("correction_tol", ctypes.c_double)]


class NewtonStats(ctypes.Structure):
    _fields_ = [("newton_iters", ctypes.c_int), ("linear_iters", ctypes.c_int),
                ("linear_failures", ctypes.c_int), ("converged", ctypes.c_int), ("final_maxres", ctypes.c_double),
                ("history", ctypes.c_double * 16), ("nhistory", ctypes.c_int),
                ("done_ms", ctypes.c_double)]


class NewtonStepState(ctypes.Structure):
    _fields_ = [("newton_iters", ctypes.c_int), ("converged", ctypes.c_int), ("failed", ctypes.c_int),
                ("maxres", ctypes.c_double), ("max_update", ctypes.c_double), ("linear_iters", ctypes.c_int),
                ("linear_failures", ctypes.c_int)]


class ArclengthState(ctypes.Structure):
    _fields_ = [("theta_squared", ctypes.c_double), ("parameter_derivative", ctypes.c_double),
                ("parameter_current", ctypes.c_double), ("ds_current", ctypes.c_double),
                ("step_taken", ctypes.c_int), ("parameter", ctypes.c_double),
                ("newton_iters", ctypes.c_int), ("rejected", ctypes.c_int),
                ("linear_failed", ctypes.c_int), ("turning_point", ctypes.c_int)]


class MultiRequest(ctypes.Structure):
    _fields_ = [("what", ctypes.c_int), ("param", ctypes.c_int), ("hvec", ctypes.c_int),
                ("transposed", ctypes.c_int), ("contribution", ctypes.c_char_p)]


# request kinds of the multi-assembly (PYHIP_MA_*), by the reference's "what" strings
MULTI_WHAT = {"residuals": (0, 0), "jacobian": (1, 0), "mass_matrix": (2, 0),
              "dresiduals_dparameter": (3, 0), "djacobian_dparameter": (4, 0),
              "dmass_matrix_dparameter": (5, 0), "hessian_vector_product": (6, 0),
              "hessian_vector_product_transposed": (6, 1),
              "mass_matrix_hessian_vector_product": (7, 0),
              "mass_matrix_hessian_vector_product_transposed": (7, 1)}


# every symbol declared in include/pyoomph_hip.h, with (restype, argtypes)
SIGNATURES = {
    "pyhip_last_error": (ctypes.c_char_p, []),
    "pyhip_version": (ctypes.c_int, []),
    "pyhip_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "pyhip_element_init": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ElementTable)]),
    "pyhip_problem_create": (ctypes.c_int, [ctypes.POINTER(SubproblemDesc), ctypes.c_int,
                                            ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "pyhip_problem_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "pyhip_problem_info": (ctypes.c_int, [ctypes.c_void_p, c_i64p, c_i64p, c_i64p, c_i64p]),
    "pyhip_problem_dof_offsets": (ctypes.c_int, [ctypes.c_void_p, c_i64p]),
    "pyhip_problem_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_dp, ctypes.c_int]),
    "pyhip_pattern_copy": (ctypes.c_int, [ctypes.c_void_p, c_i32p, c_i32p]),
    "pyhip_assemble": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "pyhip_assemble_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "pyhip_codegen_compile": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p),
                                             ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_size_t)]),
    "pyhip_codegen_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "pyhip_codegen_set_fields": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]),
    "pyhip_codegen_set_layout": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "pyhip_codegen_set_face_form": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "pyhip_codegen_set_constraint_form": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "pyhip_codegen_add_residual_table": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
    "pyhip_problem_set_solved_residual": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p,
                                                         ctypes.POINTER(ctypes.c_int)]),
    "pyhip_newton_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(NewtonOpts),
                                         ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(NewtonStepState)]),
    "pyhip_problem_set_aux": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_dp, ctypes.c_int64]),
    "pyhip_problem_set_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "pyhip_jit_element_init": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "pyhip_jit_set_form_params": (ctypes.c_int, [ctypes.c_void_p, c_dp, ctypes.c_int]),
    "pyhip_superlu": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)] * 4 + [c_dp] + [ctypes.POINTER(ctypes.c_int)] * 2 +
                      [c_dp] + [ctypes.POINTER(ctypes.c_int)] * 3 + [ctypes.POINTER(ctypes.c_void_p),
                                                                     ctypes.POINTER(ctypes.c_int)]),
    "pyhip_export_reference_csr": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_void_p, ctypes.c_void_p, c_i64p]),
    "pyhip_linear_opts_default": (None, [ctypes.POINTER(LinearOpts)]),
    "pyhip_newton_opts_default": (None, [ctypes.POINTER(NewtonOpts)]),
    "pyhip_newton_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(NewtonOpts),
                                          ctypes.POINTER(NewtonStats)]),
    "pyhip_arclength_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                            ctypes.POINTER(NewtonOpts), ctypes.POINTER(ctypes.c_int32),
                                            c_dp, ctypes.POINTER(ArclengthState), c_dp]),
    "pyhip_assemble_hessian_vector": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_void_p, ctypes.c_void_p]),
    "pyhip_set_time_stepping": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_dp, ctypes.c_void_p,
                                               ctypes.c_void_p]),
    "pyhip_multi_assembly_plan": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(MultiRequest), ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                                 ctypes.POINTER(ctypes.c_int)]),
    "pyhip_assemble_multi": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                            ctypes.POINTER(MultiRequest), ctypes.c_void_p, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
                                            ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
                                            c_i64p]),
    "pyhip_timing_enable": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "pyhip_timing_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, c_i64p,
                                         ctypes.POINTER(ctypes.c_double)]),
    "pyhip_timing_bytes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, c_i64p]),
    "pyhip_solve_serial": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          c_dp, c_i32p, c_i32p, c_dp, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_void_p)]),
}

_lib = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libpyoomph_hip.so is not built (run __graft_entry__.build() or "
                           "make -C pyoomph_amd/csrc); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != PYHIP_OK:
        msg = load().pyhip_last_error().decode(errors="replace")
        raise RuntimeError("pyoomph HIP error %d: %s" % (rc, msg))
    return rc


def ptr(t):
    """Raw pointer of a torch tensor / numpy array."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return ctypes.c_void_p(t.data_ptr())
    return t.ctypes.data_as(ctypes.c_void_p)
