"""Oracle BDF time stepping of the scalar forms (test infrastructure, like the rest of
``oracle/``).

Restates:
- the weights of MultiTimeStepper::set_weights (src/timestepper.cpp:32-59):
  BDF1 (1/dt, -1/dt, 0) and BDF2 (1/dt + 1/(dt + dtprev), -(dt + dtprev)/(dt dtprev),
  dt/((dt + dtprev) dtprev)) for the current, previous and second-previous values;
- the choice of weights for partial_t in prepare_shape_buffer_for_integration
  (src/elements.cpp:4583-4636): BDF2, degraded to BDF1 while no unsteady step has been
  completed;
- oomph's unsteady Newton step: the history values shift by one level (the previous
  values take the current ones), time advances by dt, and Newton starts from the current
  values (no predictor without adaptive time stepping). The initial condition fills every
  history level (impulsive start).
"""
import numpy as np

from .newton import newton_solve


def bdf1_weights(dt):
    return np.array([1.0 / dt, -1.0 / dt, 0.0])


def bdf2_weights(dt, dtprev):
    return np.array([1.0 / dt + 1.0 / (dt + dtprev), -(dt + dtprev) / (dt * dtprev),
                     dt / ((dt + dtprev) * dtprev)])


class TimeState:
    """Time, steps done, previous step length and the two history dof vectors."""

    def __init__(self, x0, t0=0.0):
        self.time = t0
        self.steps_done = 0
        self.dt_prev = None
        self.h1 = np.array(x0, dtype=np.float64, copy=True)
        self.h2 = self.h1.copy()

    def weights(self, dt):
        if self.steps_done == 0:
            return bdf1_weights(dt)
        return bdf2_weights(dt, self.dt_prev)


def unsteady_newton_step(problem, x, state, dt, **newton_kw):
    """One unsteady Newton step of length dt from the current values x; returns the new
    values and advances ``state``."""
    state.h2 = state.h1.copy()
    state.h1 = np.array(x, dtype=np.float64, copy=True)
    problem.set_time(state.weights(dt), state.h1, state.h2)
    try:
        x = newton_solve(problem, x, **newton_kw)
    finally:
        problem.set_time(None, None, None)
    state.steps_done += 1
    state.dt_prev = dt
    state.time += dt
    return x
