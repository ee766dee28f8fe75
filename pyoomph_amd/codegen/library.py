"""Weak forms written in the expression IR.

``poisson(ngauss)``, ``bratu()`` and ``reaction_diffusion()`` are the forms of the compiled
kernels (element_kernels.h), written so that the generated kernels evaluate the same
operations in the same order -- their FP64 results are bitwise those of the compiled
kernels. ``nonlinear_diffusion()`` has no compiled counterpart: -div((1 + k u^2) grad u) = s
with a parameter k, the kind of form a pyoomph user adds (a non-constant flux Jacobian and
a dF/du coupling).
"""
from . import expr as E
from .scalar import ScalarWeakForm


def poisson(ngauss=8):
    """weak(grad u, grad v) - weak(f, v), f = sum_k a_k exp(-|x - c_k|^2 / w_k^2);
    parameters as the compiled Poisson kernel reads them: n, then (cx, cy, w, a) per
    Gaussian (pyoomph/equations/poisson.py:70-74)."""
    def flux(s):
        return list(s.grad_u)

    def source(s):
        f = E.ZERO
        for k in range(ngauss):
            b = 1 + 4 * k
            dx = s.x[0] - s.p[b]
            dy = s.x[1] - s.p[b + 1]
            r2 = dx * dx + dy * dy
            w = s.p[b + 2]
            f = f + s.p[b + 3] * E.exp(-r2 / (w * w))
        return -f

    return ScalarWeakForm("gen_poisson%d" % ngauss, flux, source, nparams=1 + 4 * ngauss)


def bratu():
    """weak(grad u, grad v) - weak(lambda exp(u), v), global parameter lambda."""
    return ScalarWeakForm("gen_bratu", lambda s: list(s.grad_u), lambda s: -(s.p[0] * E.exp(s.u)),
                          nparams=1, global_param=0, global_param_name="lambda")


def reaction_diffusion():
    """weak(grad u, grad v) + weak(u^3 - f, v), f the auxiliary nodal field."""
    return ScalarWeakForm("gen_reaction_diffusion", lambda s: list(s.grad_u), lambda s: s.u ** 3 - s.f,
                          aux=True)


def nonlinear_diffusion():
    """weak((1 + k u^2) grad u, grad v) - weak(s0, v): parameters k, s0."""
    def flux(s):
        c = 1.0 + s.p[0] * (s.u * s.u)
        return [c * g for g in s.grad_u]

    return ScalarWeakForm("gen_nonlinear_diffusion", flux, lambda s: -s.p[1], nparams=2)


def navier_stokes():
    """Steady / transient incompressible Navier-Stokes in Taylor-Hood form, parameters
    (mu, rho): momentum flux mu (grad u + grad u^T) - p I, convection rho (u . grad) u, mass
    rho, continuity div u -- the terms of pyoomph's NavierStokesEquations
    (pyoomph/equations/navier_stokes.py:303-347) written in the hand-written kernel's order."""
    from .mixed import TaylorHoodWeakForm

    def flux(s):
        mu = s.par[0]
        F = []
        for a in range(2):
            row = []
            for d in range(2):
                v = mu * (s.grad_u[a][d] + s.grad_u[d][a])
                if a == d:
                    v = v - s.p
                row.append(v)
            F.append(row)
        return F

    def source(s):
        rho = s.par[1]
        return [rho * (s.u[0] * s.grad_u[a][0] + s.u[1] * s.grad_u[a][1]) for a in range(2)]

    def continuity(s):
        return s.grad_u[0][0] + s.grad_u[1][1]

    return TaylorHoodWeakForm("gen_navier_stokes", flux, source, continuity, mass=lambda s: s.par[1], nparams=2)


def stokes_brinkman():
    """A Taylor-Hood form with no hand-written counterpart: Stokes flow with a Brinkman drag,
    mu (grad u + grad u^T) - p I, drag k u, body force (0, -g), parameters (mu, k, g)."""
    from .mixed import TaylorHoodWeakForm

    def flux(s):
        mu = s.par[0]
        return [[mu * (s.grad_u[a][d] + s.grad_u[d][a]) - (s.p if a == d else 0.0) for d in range(2)]
                for a in range(2)]

    def source(s):
        return [s.par[1] * s.u[0], s.par[1] * s.u[1] + s.par[2]]

    def continuity(s):
        return s.grad_u[0][0] + s.grad_u[1][1]

    return TaylorHoodWeakForm("gen_stokes_brinkman", flux, source, continuity, nparams=3)
