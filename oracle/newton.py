"""Oracle Newton solver with the reference's SuperLU call pattern.

Test infrastructure only (see ``oracle/__init__``).

Restates oomph ``Problem::newton_solve`` (problem.cc:8986-9345) as driven by pyoomph
(``steady_newton_solve`` problem.cc:9495-9577; Python defaults
pyoomph/generic/problem.py:409-411,536: tolerance 1e-8, at most 10 iterations,
``max_residuals`` 1e10, always take one Newton step; relaxation 1.0 problem.cc:96):

    count = 1: residuals -> maxres (max |r|, any non-finite entry -> inf,
               double_vector.cc:611-630); stop early only if maxres < tol and not
               always_take_one_newton_step
    each step: (r, J) = get_jacobian; solve J dx = r; x -= dx;
               r = get_residuals; maxres; stop if maxres < tol;
               raise NewtonSolverError if maxres > max_residuals or count == max_iter.

The linear solve is SuperLU the way pyoomph calls it: oomph hands over the CSR arrays as
a compressed-column matrix (that is J^T, linear_solver.cc:2253-2311) and
pyoomph/solvers/scipy.py:67,96 runs ``splu(csc_matrix((values, rowind, colptr)))``
followed by ``lu.solve(b, "T")``, which solves J x = b.

Parity: pinned by tests/test_oracle_newton.py (converged solution against analytic /
manufactured solutions, quadratic convergence, error paths).
"""
import numpy as np
import scipy.sparse
import scipy.sparse.linalg


class NewtonSolverError(RuntimeError):
    def __init__(self, count, maxres):
        super().__init__("Newton solver failed after %d steps, max residual %g" % (count, maxres))
        self.count, self.maxres = count, maxres


def max_abs(v):
    """DoubleVector::max with pyoomph's NaN guard (double_vector.cc:611-630)."""
    v = np.asarray(v)
    if v.size == 0:
        return 0.0
    if not np.all(np.isfinite(v)):
        return float("inf")
    return float(np.max(np.abs(v)))


def superlu_solve(n, row_start, col, val, rhs):
    """scipy.py:67 + :96 with transpose=1: splu of the CSR read as CSC, solve with "T".

    The arrays are copied: scipy sorts a CSC matrix's indices in place, which would
    reorder the caller's values relative to its column indices."""
    A = scipy.sparse.csc_matrix((np.array(val, dtype=np.float64, copy=True),
                                 np.array(col, dtype=np.int32, copy=True),
                                 np.array(row_start, dtype=np.int32, copy=True)), shape=(n, n))
    lu = scipy.sparse.linalg.splu(A)
    return lu.solve(np.asarray(rhs, dtype=np.float64), "T")


def newton_solve(problem, x0, tol=1e-8, max_iter=10, max_residuals=1e10,
                 always_take_one_newton_step=True, history=None):
    x = np.array(x0, dtype=np.float64, copy=True)
    n = problem.ndof
    count = 0
    while True:
        count += 1
        if count == 1:
            r = problem.get_residuals(x)
            maxres = max_abs(r)
            if history is not None:
                history.append(maxres)
            if maxres < tol and not always_take_one_newton_step:
                break
        r, rs, col, val = problem.get_jacobian(x)
        dx = superlu_solve(n, rs, col, val, r)
        x -= 1.0 * dx
        r = problem.get_residuals(x)
        maxres = max_abs(r)
        if history is not None:
            history.append(maxres)
        if maxres < tol:
            break
        if maxres > max_residuals or count == max_iter:
            raise NewtonSolverError(count, maxres)
    return x


def converged_solve(problem, x0, tol=1e-12, max_iter=30):
    """The accuracy reference: ``newton_solve`` to max|R| < tol, then one more Newton step.

    Not a restatement of the reference: an absolute residual test stops an oracle solve at
    an error of up to (tol / initial residual) relative, which on a fine 3D mesh (small
    residual scale) is not below 1e-8. The extra step of a quadratically converging Newton
    iteration takes the error to roundoff (the next residual is ~ the square of the last)."""
    x = newton_solve(problem, x0, tol=tol, max_iter=max_iter)
    return newton_solve(problem, x, tol=tol, max_iter=2)
