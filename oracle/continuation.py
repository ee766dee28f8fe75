"""Oracle pseudo-arclength continuation in the Bratu lambda (test infrastructure, like the
rest of ``oracle/``).

Restates oomph's Problem::arc_length_step_solve (thirdparty/oomph-lib problem.cc
10718-11080; no spatial adaptation, no bifurcation detection, no continuation time
stepper, analytic derivatives) with newton_solve_continuation (:9599-9917: block
elimination of the bordered system, J y = R and J z = dR/dlambda, both with SuperLU),
calculate_continuation_derivatives (:9918-10020, :10115-10200; Scale_arc_length with
Desired_proportion_of_arc_length 0.5) and the step-length rule (Desired_newton_iterations_ds
5; Minimum_ds 1e-10; a failed Newton solve retries with ds * 2/3). Defaults of the state:
Theta_squared 1, Parameter_derivative 1, Dof_derivative 0 (problem.cc:115-140).

turning_point: the step's Parameter_derivative has the opposite sign of the previous step's
(a fold of the branch between the two parameter values). The bordered matrix
[[J, dR/dlambda], [Theta^2 Dof_derivative^T, Parameter_derivative]] has the determinant
det J * (Parameter_derivative - Theta^2 Dof_derivative . z), J z = dR/dlambda, and keeps
its sign along a branch without branch points, while the new Parameter_derivative takes the
sign of the continuation direction Parameter_derivative - Dof_derivative . z (:10138-10183):
its sign change marks the fold, where det J changes sign and oomph's Sign_of_jacobian
check (problem.cc:10944-10975, off unless Bifurcation_detection is set) would report it.
"""
import numpy as np

from .newton import NewtonSolverError, max_abs, superlu_solve


class ArclengthState:
    def __init__(self, ndof):
        self.theta_squared = 1.0
        self.parameter_derivative = 1.0
        self.parameter_current = 0.0
        self.ds_current = 0.0
        self.step_taken = False
        self.dof_current = np.zeros(ndof)
        self.dof_derivative = np.zeros(ndof)
        self.newton_iters = 0
        self.turning_point = False


def _get_lam(op):
    return float(op.params[0])


def _set_lam(op, lam):
    op.params = np.array([lam], dtype=np.float64)


def _constraint(op, x, st):
    c = float(np.dot(st.dof_derivative, x - st.dof_current))
    c *= st.theta_squared
    c += st.parameter_derivative * (_get_lam(op) - st.parameter_current) - st.ds_current
    return c


def newton_solve_continuation(op, x, st, tol=1e-8, max_iter=10, max_residuals=1e10):
    """Returns (count, x, z); raises NewtonSolverError like oomph."""
    n = op.ndof
    count = 0
    r = op.get_residuals(x)
    acr = _constraint(op, x, st)      # count == 1 check (always one Newton step)
    z = None
    while True:
        count += 1
        res, rs, col, val = op.get_jacobian(x)
        y = superlu_solve(n, rs, col, val, res)
        dr = op.get_residuals(x, deriv=1)
        z = superlu_solve(n, rs, col, val, dr)
        uy = float(np.dot(st.dof_derivative, y)) * st.theta_squared
        uz = float(np.dot(st.dof_derivative, z)) * st.theta_squared
        dparam = (acr - uy) / (st.parameter_derivative - uz)
        _set_lam(op, _get_lam(op) - dparam)
        x = x - (y - dparam * z)
        r = op.get_residuals(x)
        maxres = max_abs(r)
        acr = _constraint(op, x, st)
        maxres = max(maxres, abs(acr))
        if maxres < tol:
            return count, x, z
        if maxres > max_residuals or count == max_iter:
            raise NewtonSolverError(count, maxres)


def _helper(st, z, der_dot_z):
    cdir = st.parameter_derivative - der_dot_z
    pd = 1.0 / np.sqrt(1.0 + st.theta_squared * float(np.dot(z, z)))
    if pd * cdir < 0.0:
        pd = -pd
    st.parameter_derivative = pd
    st.dof_derivative = -pd * z


def arc_length_step(op, x, st, ds, desired_newton_iterations=5, min_ds=1e-10, **newton_kw):
    """One step; returns (x, next ds). The Bratu lambda is op.params[0]."""
    st.parameter_current = _get_lam(op)
    st.dof_current = np.array(x, dtype=np.float64, copy=True)
    st.ds_current = ds
    while True:
        if abs(st.ds_current) < min_ds:
            raise RuntimeError("DESIRED ARC-LENGTH STEP HAS FALLEN BELOW MINIMUM TOLERANCE")
        lam = st.parameter_current + st.parameter_derivative * st.ds_current
        _set_lam(op, lam)
        st.ds_current = (lam - st.parameter_current) / st.parameter_derivative
        x = st.dof_current + st.dof_derivative * st.ds_current
        try:
            count, x, z = newton_solve_continuation(op, x, st, **newton_kw)
            break
        except NewtonSolverError:
            st.ds_current *= 2.0 / 3.0
    st.newton_iters = count
    pd_before, taken_before = st.parameter_derivative, st.step_taken
    _helper(st, z, float(np.dot(st.dof_derivative, z)))
    pd = st.parameter_derivative
    if pd != 1.0:
        st.theta_squared *= (pd * pd / 0.5) * ((1.0 - 0.5) / (1.0 - pd * pd))
        _helper(st, z, float(np.dot(st.dof_derivative, z)))
    if not st.step_taken:
        st.ds_current = (_get_lam(op) - st.parameter_current) / st.parameter_derivative
    st.turning_point = bool(taken_before) and pd_before * st.parameter_derivative < 0.0
    st.step_taken = True
    if count < desired_newton_iterations:
        return x, st.ds_current * 1.5
    if count > desired_newton_iterations:
        return x, st.ds_current * (2.0 / 3.0)
    return x, st.ds_current
