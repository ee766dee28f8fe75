"""Weak forms on the hot path, as equation objects with pyoomph-like constructors.

Each maps to one batched device kernel family of libpyoomph_hip.so (the functions the
reference's code generator emits for the same weak form, src/codegen.cpp:3912-4109):

* ``PoissonEquation``      -- weak(grad u, grad v) - weak(f, v)
                              (pyoomph/equations/poisson.py:70-74); the source is a sum of
                              Gaussians a*exp(-|x-c|^2/w^2) evaluated at integration points.
* ``BratuEquation``        -- weak(grad u, grad v) - weak(lambda*exp(u), v), global
                              parameter ``lambda`` with dR/dlambda.
* ``ReactionDiffusionEquation`` -- weak(grad u, grad v) + weak(u^3 - f, v), f a nodal field.
* ``NavierStokesEquations`` -- steady incompressible Navier-Stokes, Taylor-Hood
                              (pyoomph/equations/navier_stokes.py:255-347, 481-503):
                              weak(2 mu sym(grad u) - p I, grad v) + weak(rho u.grad u, v)
                              + weak(div u, q).
"""
import numpy as np

from . import _lib

__all__ = ["Equations", "GaussianSource", "PoissonEquation", "BratuEquation",
           "ReactionDiffusionEquation", "NavierStokesEquations"]

_SPACE_KINDS = {"C1": ("tri_p1",), "C2": ("tri_p2", "quad_q2", "hex_q2")}


class Equations:
    form = None
    nvalue = 1
    space = "C2"
    symmetric = False   # Jacobian symmetric positive definite (conjugate gradients apply)

    def params(self):
        return np.zeros(0)

    def aux(self, mesh):
        return None

    def dummy_values(self, mesh):
        return None

    def check_mesh(self, mesh):
        if mesh.kind not in _SPACE_KINDS[self.space]:
            raise ValueError("mesh of kind %s does not match space %s" % (mesh.kind, self.space))


class GaussianSource:
    """f(x) = sum_k a_k exp(-|x - c_k|^2 / w_k^2) (2D)."""

    def __init__(self, centres, widths, amplitudes):
        self.centres = np.asarray(centres, dtype=np.float64).reshape(-1, 2)
        self.widths = np.asarray(widths, dtype=np.float64).reshape(-1)
        self.amplitudes = np.asarray(amplitudes, dtype=np.float64).reshape(-1)
        if not (len(self.centres) == len(self.widths) == len(self.amplitudes)):
            raise ValueError("inconsistent Gaussian source")
        if len(self.widths) > 15:
            raise ValueError("at most 15 Gaussians")

    def flat(self):
        """(cx, cy, w, a) per Gaussian."""
        return np.stack([self.centres[:, 0], self.centres[:, 1], self.widths, self.amplitudes],
                        axis=1).reshape(-1)

    def __call__(self, x, y):
        f = np.zeros_like(np.asarray(x, dtype=np.float64))
        for (cx, cy), w, a in zip(self.centres, self.widths, self.amplitudes):
            f = f + a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (w * w))
        return f


class PoissonEquation(Equations):
    form = _lib.FORM_POISSON
    symmetric = True

    def __init__(self, name="u", *, space="C2", source=None, coefficient=1):
        if coefficient != 1:
            raise ValueError("only coefficient=1 has a device kernel")
        if source is not None and not isinstance(source, GaussianSource):
            raise TypeError("source must be a GaussianSource (or None)")
        self.name, self.space, self.source = name, space, source

    def params(self):
        if self.source is None:
            return np.zeros(1)
        return np.concatenate([[len(self.source.widths)], self.source.flat()])


class BratuEquation(Equations):
    form = _lib.FORM_BRATU
    symmetric = True   # lambda below the fold of the solution branch
    global_parameters = ("lambda",)

    def __init__(self, name="u", *, space="C2", lambda_=1.0):
        self.name, self.space, self.lambda_ = name, space, float(lambda_)

    def params(self):
        return np.array([self.lambda_])


class ReactionDiffusionEquation(Equations):
    form = _lib.FORM_REACTION_DIFFUSION
    symmetric = True

    def __init__(self, name="u", *, space="C2", source_field=None):
        if source_field is None:
            raise ValueError("source_field (nodal values of f) is required")
        self.name, self.space = name, space
        self.source_field = np.asarray(source_field, dtype=np.float64)

    def aux(self, mesh):
        if self.source_field.shape != (mesh.nnode,):
            raise ValueError("source_field must have one value per node")
        return self.source_field


class NavierStokesEquations(Equations):
    """Values per node: [velocity_x, velocity_y, pressure]; pressure lives on the
    element vertices (C1), the other nodes carry a pinned dummy."""
    form = _lib.FORM_NAVIER_STOKES_TH
    nvalue = 3

    def __init__(self, *, dynamic_viscosity=1.0, mass_density=1.0, mode="TH"):
        if mode != "TH":
            raise ValueError("only Taylor-Hood (mode='TH') has a device kernel")
        self.mu, self.rho = float(dynamic_viscosity), float(mass_density)
        self.space = "C2"

    def params(self):
        return np.array([self.mu, self.rho])

    def check_mesh(self, mesh):
        if mesh.kind != "quad_q2":
            raise ValueError("Taylor-Hood needs a Q2 quadrilateral mesh")

    def dummy_values(self, mesh):
        vertex = np.zeros(mesh.nnode, dtype=bool)
        vertex[mesh.elements[:, [0, 2, 6, 8]].reshape(-1)] = True
        d = np.zeros((mesh.nnode, 3), dtype=bool)
        d[~vertex, 2] = True
        return d
