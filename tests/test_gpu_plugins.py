"""The pyoomph plugin classes (pyoomph_amd/plugins.py) driven the way pyoomph drives its
plugins, against the oracle.

The host is a stand-in with the two members of pyoomph's Problem that the plugins read
(``get_current_dofs()`` returning (dofs, flags), ``get_global_parameter(name).value``);
the Newton loop is oomph's (problem.cc:9100-9300: residual, J d = R, x -= d, max-abs
residual test), taking R and J from ``HIPAssembler.get_residuals_and_jacobian`` (as
pyoomph's C++ core does when a custom assembler is set, src/problem.cpp:944-968) and the
solve from ``HIPKrylovSerial.solve_serial`` with the arguments oomph's SuperLU driver
passes (compressed columns of J^T, transpose 1; linear_solver.cc:2253-2311). The solution
must be within 1e-8 of the oracle's SuperLU Newton.
"""
import numpy as np
import pytest

from oracle import newton as on
from pyoomph_amd import workloads as wl
from tests.oracle_bridge import oracle_problem

pytestmark = pytest.mark.gpu


class _Param:
    def __init__(self, v):
        self.value = v


class HostProblem:
    """The members of pyoomph's Problem the plugins use."""

    def __init__(self, ndof, params):
        self.x = np.zeros(ndof)
        self.params = {k: _Param(v) for k, v in params.items()}

    def get_current_dofs(self):
        return self.x.copy(), [False] * len(self.x)

    def get_global_parameter(self, name):
        return self.params[name]


@pytest.mark.parametrize("space,lam", [("C2", 2.2), ("C1", 0.8)])
def test_plugins_drive_an_oomph_newton_loop(gpu, space, lam):
    from pyoomph_amd.plugins import HIPAssembler, HIPKrylovSerial
    N = 12
    host = HostProblem(0, {"lambda": lam})
    asm = HIPAssembler(wl.bratu(N, space, 0.0), parameters=["lambda"])
    host.x = np.zeros(asm.hip.ndof)
    asm._set_problem(host)
    asm.initialize()
    solver = HIPKrylovSerial(host)
    n = asm.hip.ndof
    count = 0
    R = asm.get_residuals_and_jacobian(False)
    while True:
        count += 1
        R, J = asm.get_residuals_and_jacobian(True)
        assert J.indices.dtype == np.int32 and J.data.dtype == np.float64
        # oomph hands its CSR over as the CSC arrays of J^T with transpose = 1
        assert solver.solve_serial(1, n, J.nnz, 0, J.data, J.indices, J.indptr, None, n, 1) == 0
        d = R.copy()
        assert solver.solve_serial(2, n, 0, 1, None, None, None, d, n, 1) == 0
        host.x = host.x - d
        R = asm.get_residuals_and_jacobian(False)
        if np.max(np.abs(R)) < 1e-8 or count >= 10:
            break
    assert np.max(np.abs(R)) < 1e-8
    op = oracle_problem("bratu", N=N, space=space, lam=lam)
    xo = on.newton_solve(op, np.zeros(op.ndof))
    assert np.max(np.abs(host.x - xo)) <= 1e-8 * np.max(np.abs(xo))
    # dR/dlambda through the assembler, like get_derivative_wrt_global_parameter
    dR = asm.get_residuals_and_jacobian(False, "lambda")
    assert np.max(np.abs(dR - op.get_residuals(host.x, deriv=1))) <= 1e-12 * np.max(np.abs(dR))
    solver.solve_serial(3, n, 0, 0, None, None, None, None, n, 1)
