"""pyoomph plugin classes that route the hot path through libpyoomph_hip.so.

* ``HIPKrylovSerial`` -- a linear-solver plugin with the interface of
  ``GenericLinearSystemSolver`` (pyoomph/solvers/generic.py:60-95): ``solve_serial(op_flag,
  n, nnz, nrhs, values, rowind, colptr, b, ldb, transpose)`` with op 1 = factorise,
  2 = solve in place in ``b``, 3 = free, returning 0 (the sign-of-determinant slot, which
  the reference's SuperLU plugin also leaves at 0, pyoomph/solvers/scipy.py:103).
* ``HIPAssembler`` -- a custom assembler with the interface of ``CustomAssemblyBase``
  (pyoomph/generic/assembly.py:36-89): pyoomph's C++ core calls
  ``get_residuals_and_jacobian(require_jacobian, dparameter)`` on every residual /
  Jacobian request (src/problem.cpp:944-968) and receives the residual and, if asked, the
  CSR Jacobian in the reference's form (int32 / float64, first-touch columns, exact zeros
  dropped). The current dofs come from ``self.problem.get_current_dofs()[0]`` and the
  global parameters from ``self.problem.get_global_parameter(name).value``, as the
  reference's own assemblers read them (pyoomph/generic/assembly.py:161-180).

Where pyoomph is importable the classes derive from its base classes (and the solver is
registered under ``"hip"``); otherwise from local bases with the same interface, so the
classes can be used and tested without pyoomph.
"""
import ctypes

import numpy as np

from . import _lib

try:  # pragma: no cover - pyoomph is not installed in the build environment
    from pyoomph.generic.assembly import CustomAssemblyBase as _AssemblyBase
    from pyoomph.solvers.generic import GenericLinearSystemSolver as _SolverBase
    _HAVE_PYOOMPH = True
except ImportError:
    _HAVE_PYOOMPH = False

    class _SolverBase:
        """The interface of pyoomph's GenericLinearSystemSolver that the plugin uses."""

        def __init__(self, problem):
            self.problem = problem

        def setup_solver(self):
            pass

        def distributed_possible(self):
            return True

    class _AssemblyBase:
        """The interface of pyoomph's CustomAssemblyBase that the plugin uses."""

        def __init__(self):
            self.problem = None

        def _set_problem(self, problem):
            self.problem = problem

        def has_custom_solve_routine(self):
            return False

        def invalidate_cache(self):
            pass

        def initialize(self):
            pass

        def finalize(self):
            pass


class HIPKrylovSerial(_SolverBase):
    """Device Krylov solve behind pyoomph's serial solver interface."""
    idname = "hip"

    def __init__(self, problem):
        super().__init__(problem)
        self._handle = ctypes.c_void_p()

    def solve_serial(self, op_flag, n, nnz, nrhs, values, rowind, colptr, b, ldb, transpose):
        L = _lib.load()
        keep = []

        def arr(a, dtype, ctype):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return a.ctypes.data_as(ctype)

        bp = None
        if b is not None:
            if not (isinstance(b, np.ndarray) and b.dtype == np.float64 and b.flags.c_contiguous):
                raise TypeError("b must be a contiguous float64 array (solved in place)")
            bp = b.ctypes.data_as(_lib.c_dp)
        rc = L.pyhip_solve_serial(int(op_flag), int(n), int(nnz), int(nrhs), arr(values, np.float64, _lib.c_dp),
                                  arr(rowind, np.int32, _lib.c_i32p), arr(colptr, np.int32, _lib.c_i32p), bp,
                                  int(ldb) if ldb else int(n), int(transpose), ctypes.byref(self._handle))
        _lib.check(rc)
        return 0

    def distributed_possible(self):
        return False

    def __del__(self):
        if getattr(self, "_handle", None) is not None and self._handle.value:
            try:
                _lib.load().pyhip_solve_serial(3, 0, 0, 0, None, None, None, None, 0, 0, ctypes.byref(self._handle))
            except Exception:
                pass


if _HAVE_PYOOMPH:  # pragma: no cover
    HIPKrylovSerial = _SolverBase.register_solver()(HIPKrylovSerial)


class HIPAssembler(_AssemblyBase):
    """Residual and Jacobian of the host problem, assembled on the device.

    ``mesh_problem``: the pyoomph_amd ``MeshProblem`` describing the host problem's mesh,
    equations and Dirichlet conditions with the same node and equation numbering
    (pyoomph_amd.meshes reproduces RectangularQuadMesh / CuboidBrickMesh numbering).
    ``parameters``: names of the host's global parameters the form reads (Bratu:
    ["lambda"]), read from the host before every assembly.
    """

    def __init__(self, mesh_problem, parameters=(), device="cuda:0"):
        super().__init__()
        from .problem import Problem
        self.hip = Problem(mesh_problem, device=device)
        self.parameters = list(parameters)
        if len(self.parameters) > 1 or any(n != "lambda" for n in self.parameters):
            raise ValueError("the device forms have one global parameter, the Bratu 'lambda'")

    def _sync_from_host(self):
        if self.problem is None:
            raise RuntimeError("the assembler is not attached to a problem (set_custom_assembler)")
        dofs = np.asarray(self.problem.get_current_dofs()[0], dtype=np.float64)
        if dofs.shape != (self.hip.ndof,):
            raise ValueError("host problem has %d dofs, the device problem %d" % (dofs.size, self.hip.ndof))
        self.hip.set_current_dofs(dofs)
        for name in self.parameters:
            self.hip.set_global_parameter(0, float(self.problem.get_global_parameter(name).value))

    def get_residuals_and_jacobian(self, require_jacobian, dparameter=None):
        self._sync_from_host()
        dp = None if dparameter in (None, "") else dparameter
        if not require_jacobian:
            return self.hip.get_residuals(dparameter=dp)
        return self.hip.assemble_jacobian(dparameter=dp)
