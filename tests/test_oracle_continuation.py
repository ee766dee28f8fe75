"""Pins the oracle's pseudo-arclength continuation (no GPU): on the 2D Bratu problem
-lap u = lambda e^u on the unit square the branch from u = 0 turns at the fold
lambda_c = 6.808 (the known value of the continuous problem: Mohsen, 'A simple solution of
the Bratu problem', Comput. Math. Appl. 67 (2014), 26-33, lambda_c = 6.80812); the
continuation has to pass it and come back down the upper branch, every point being a
solution (max|R| < 1e-8) and the arclength constraint holding."""
import numpy as np

from oracle import continuation as oc
from oracle.newton import max_abs
from tests.oracle_bridge import oracle_problem


def _run(N, space, steps, ds=0.5, max_ds=0.3):
    """Steps with pyoomph's max_ds clamp (pyoomph/generic/problem.py:3578-3580)."""
    op = oracle_problem("bratu", N=N, space=space, lam=0.0)
    st = oc.ArclengthState(op.ndof)
    x = np.zeros(op.ndof)
    lams, norms = [0.0], [0.0]
    for _ in range(steps):
        ds = min(abs(ds), max_ds) * (1 if ds > 0 else -1)
        x, ds = oc.arc_length_step(op, x, st, ds)
        assert max_abs(op.get_residuals(x)) < 1e-8
        lams.append(float(op.params[0]))
        norms.append(float(np.max(np.abs(x))))
    return np.array(lams), np.array(norms)


def test_bratu_branch_turns_at_the_fold():
    lams, norms = _run(8, "C2", 45)
    k = int(np.argmax(lams))
    assert 0 < k < len(lams) - 1, "the continuation did not pass the fold"
    # lambda as a quadratic in max u through the three points around the turn: its vertex
    c = np.polyfit(norms[k - 1:k + 2], lams[k - 1:k + 2], 2)
    lam_fold = c[2] - c[1] ** 2 / (4 * c[0])
    assert abs(lam_fold - 6.808) < 0.02      # P2 on 8 x 8 cells, three-point fit
    assert np.all(np.diff(norms) > 0)          # max u grows monotonically along the branch
    assert lams[-1] < lams[k] - 0.5            # back down the upper branch


def test_lower_branch_points_are_steady_solutions():
    from oracle.newton import newton_solve
    lams, norms = _run(6, "C1", 5)
    for lam, nm in zip(lams[1:], norms[1:]):
        op = oracle_problem("bratu", N=6, space="C1", lam=lam)
        x = newton_solve(op, np.zeros(op.ndof))
        assert abs(np.max(np.abs(x)) - nm) < 1e-8 * max(1.0, nm)


def test_turning_point_flag_marks_the_sign_change_of_det_j():
    """The turning-point flag against the sign of det J (dense LU, numpy.linalg.slogdet) at
    every continuation point: set on exactly the steps across which det J changes sign, which
    for Bratu is the one step over the fold (lambda grows up to the step's start and falls
    from its end on)."""
    op = oracle_problem("bratu", N=6, space="C1", lam=0.0)
    st = oc.ArclengthState(op.ndof)
    x = np.zeros(op.ndof)
    ds, flags, signs, lams = 0.5, [], [], [0.0]
    _, rs, col, val = op.get_jacobian(x)
    signs.append(np.linalg.slogdet(_dense(op.ndof, rs, col, val))[0])
    for _ in range(40):
        ds = min(abs(ds), 0.3) * (1 if ds > 0 else -1)
        x, ds = oc.arc_length_step(op, x, st, ds)
        flags.append(st.turning_point)
        lams.append(float(op.params[0]))
        _, rs, col, val = op.get_jacobian(x)
        signs.append(np.linalg.slogdet(_dense(op.ndof, rs, col, val))[0])
    signs = np.array(signs)
    assert np.all(signs != 0)
    changes = list(signs[1:] != signs[:-1])
    assert changes == flags
    assert sum(flags) == 1
    k = flags.index(True)              # the step from lams[k] to lams[k + 1]
    d = np.diff(lams)
    assert np.all(d[:k] > 0) and np.all(d[k + 1:] < 0)


def _dense(n, rs, col, val):
    A = np.zeros((n, n))
    for i in range(n):
        for q in range(rs[i], rs[i + 1]):
            A[i, col[q]] += val[q]
    return A
