"""Seeded synthetic workloads of the assembly-and-solve benchmarks.

* ``poisson_gaussians``   -- 2D Poisson on the unit square, ``RectangularQuadMesh(N,
  split_in_tris="left")`` with P2 triangles, f = sum of 8 Gaussians with centres U[0,1]^2,
  widths U(0.05, 0.2) and amplitudes N(0, 1), u = 0 on the boundary (64 x 64: 16,129 dofs).
* ``lid_driven_cavity``   -- steady Navier-Stokes, Taylor-Hood Q2/Q1 on N x N quads,
  lid velocity (1, 0) on "top", no-slip on the other sides (which win at the two top
  corners), pressure pinned to 0 at node 0; Re = 100 via rho = 1, mu = 1/Re
  (256 x 256: 592,387 unknowns).
* ``reaction_diffusion_cube`` -- -lap u + u^3 = f on the unit cube, Q2 bricks N^3, u = 0 on
  the boundary, f ~ U(-1, 1) per node smoothed by 3 Jacobi sweeps on the nodal lattice
  graph (each node averaged with its axis neighbours) (64^3: 2,146,689 nodes).
* ``bratu_sweep``         -- independent -lap u = lambda e^u problems on the unit square,
  N ~ U{16..128}, P1 or P2 chosen 50/50, lambda ~ U(0.5, 3.0), u = 0 on the boundary.
"""
import numpy as np

from .equations import (BratuEquation, GaussianSource, NavierStokesEquations,
                        PoissonEquation, ReactionDiffusionEquation)
from .meshes import CuboidBrickMesh, RectangularQuadMesh
from .problem import MeshProblem

ALL2D = ("left", "right", "bottom", "top")
ALL3D = ("left", "right", "bottom", "top", "back", "front")


def gaussian_source(seed, n=8):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1.0, size=(n, 2))
    w = rng.uniform(0.05, 0.2, size=n)
    a = rng.normal(0.0, 1.0, size=n)
    return GaussianSource(c, w, a)


def poisson_gaussians(N=64, seed=0):
    mesh = RectangularQuadMesh(N, split_in_tris="left", space="C2")
    eq = PoissonEquation("u", space="C2", source=gaussian_source(seed))
    return MeshProblem(mesh, eq, dirichlet=[(ALL2D, {0: 0.0})])


def lid_driven_cavity(N=256, Re=100.0):
    mesh = RectangularQuadMesh(N, split_in_tris=False, space="C2")
    eq = NavierStokesEquations(dynamic_viscosity=1.0 / Re, mass_density=1.0)
    bcs = [(("top",), {0: 1.0, 1: 0.0}),
           (("left", "right", "bottom"), {0: 0.0, 1: 0.0})]
    return MeshProblem(mesh, eq, dirichlet=bcs, pin_extra=[(0, 2, 0.0)])


def smoothed_noise(mesh, seed, sweeps=3):
    """U(-1,1) per node, then Jacobi averaging with the axis neighbours on the lattice."""
    rng = np.random.default_rng(seed)
    f = rng.uniform(-1.0, 1.0, size=mesh.nnode)
    L = 2 * mesh.N + 1
    lat = mesh.lattice
    grid = np.zeros((L,) * mesh.dim)
    idx = tuple(lat[:, d] for d in range(mesh.dim))
    grid[idx] = f
    for _ in range(sweeps):
        s = grid.copy()
        cnt = np.ones_like(grid)
        for d in range(mesh.dim):
            lo = [slice(None)] * mesh.dim
            hi = [slice(None)] * mesh.dim
            lo[d] = slice(0, L - 1)
            hi[d] = slice(1, L)
            s[tuple(lo)] += grid[tuple(hi)]
            cnt[tuple(lo)] += 1
            s[tuple(hi)] += grid[tuple(lo)]
            cnt[tuple(hi)] += 1
        grid = s / cnt
    return grid[idx]


def reaction_diffusion_cube(N=64, seed=0):
    mesh = CuboidBrickMesh(N)
    f = smoothed_noise(mesh, seed)
    eq = ReactionDiffusionEquation("u", space="C2", source_field=f)
    return MeshProblem(mesh, eq, dirichlet=[(ALL3D, {0: 0.0})])


def bratu(N, space, lam):
    mesh = RectangularQuadMesh(N, split_in_tris="left", space=space)
    eq = BratuEquation("u", space=space, lambda_=lam)
    return MeshProblem(mesh, eq, dirichlet=[(ALL2D, {0: 0.0})])


def bratu_flux(N, space, lam, alpha, beta, flux_sides=("right", "top"), fixed_sides=("left", "bottom")):
    """Bratu with u = 0 on ``fixed_sides`` and the boundary flux alpha*u + beta
    (``NeumannBC(u=alpha*var("u") + beta) @ side``) on ``flux_sides``."""
    mesh = RectangularQuadMesh(N, split_in_tris="left", space=space)
    eq = BratuEquation("u", space=space, lambda_=lam)
    return MeshProblem(mesh, eq, dirichlet=[(tuple(fixed_sides), {0: 0.0})],
                       flux=[(tuple(flux_sides), alpha, beta)])


def bratu_sweep_specs(nprob=512, seed=0, nmin=16, nmax=128):
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(nprob):
        N = int(rng.integers(nmin, nmax + 1))
        space = "C1" if rng.random() < 0.5 else "C2"
        lam = float(rng.uniform(0.5, 3.0))
        specs.append((N, space, lam))
    return specs


def bratu_sweep(nprob=512, seed=0, nmin=16, nmax=128):
    return [bratu(N, sp, lam) for N, sp, lam in bratu_sweep_specs(nprob, seed, nmin, nmax)]
