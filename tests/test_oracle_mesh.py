"""Pins the oracle's mesh numbering (hand-counted tiny meshes, closed-form counts) and
checks that the product's vectorised meshes reproduce it bit for bit."""
import numpy as np
import pytest

from oracle import meshes as om
from pyoomph_amd import meshes as pm
from pyoomph_amd import workloads as wl
from tests.oracle_bridge import oracle_problem, product_problem


def test_p2_single_cell_hand_numbering():
    m = om.rectangular_quad_mesh(1, "left", "C2")
    # vertices requested n00, n10, n01, n11; then triangle (0,1,3) mids (0,1),(1,3),(0,3);
    # triangle (0,3,2) mids (0,3) [exists], (3,2), (0,2)
    assert m.elements.tolist() == [[0, 1, 3, 4, 5, 6], [0, 3, 2, 6, 7, 8]]
    assert m.coords.tolist() == [[0, 0], [1, 0], [0, 1], [1, 1], [0.5, 0], [1, 0.5],
                                 [0.5, 0.5], [0.5, 1], [0, 0.5]]


def test_vertex_request_order_two_columns():
    m = om.rectangular_quad_mesh(2, "left", "C1")
    # ix = 0: (0,0)(1,0)(0,1)(1,1) then iy=1 adds (0,2)(1,2); ix = 1 adds (2,0),(2,1),(2,2)
    expect = [(0, 0), (1, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 0), (2, 1), (2, 2)]
    assert [tuple(c) for c in (m.coords * 2).round().astype(int).tolist()] == expect


def test_q2_cell_node_order():
    m = om.rectangular_quad_mesh(1, False, "C2")
    # lexicographic Q2 order of oomph: s0 fastest
    xy = m.coords[m.elements[0]]
    assert xy.tolist() == [[0, 0], [0.5, 0], [1, 0], [0, 0.5], [0.5, 0.5], [1, 0.5],
                           [0, 1], [0.5, 1], [1, 1]]


def test_brick_node_order():
    m = om.cuboid_brick_mesh(1)
    xyz = m.coords[m.elements[0]]
    expect = [[x, y, z] for z in (0, 0.5, 1) for y in (0, 0.5, 1) for x in (0, 0.5, 1)]
    assert xyz.tolist() == expect
    # creation order of the intermediate nodes (meshtemplate.cpp:579-621)
    first_mid = [tuple(c) for c in m.coords[8:13].tolist()]
    assert first_mid == [(0.5, 0, 0), (0, 0.5, 0), (0.5, 0.5, 0), (1, 0.5, 0), (0.5, 1, 0)]


@pytest.mark.parametrize("N", [1, 2, 5])
def test_counts(N):
    assert om.rectangular_quad_mesh(N, "left", "C2").nnode == (2 * N + 1) ** 2
    assert om.rectangular_quad_mesh(N, "left", "C2").nelement == 2 * N * N
    assert om.rectangular_quad_mesh(N, "left", "C1").nnode == (N + 1) ** 2
    assert om.rectangular_quad_mesh(N, False, "C2").nelement == N * N
    m3 = om.cuboid_brick_mesh(N)
    assert m3.nnode == (2 * N + 1) ** 3 and m3.nelement == N ** 3
    # boundary set by intersection inheritance equals the geometric boundary
    on = m3.on_boundary()
    geo = np.any((m3.coords == 0) | (m3.coords == 1), axis=1)
    assert np.array_equal(on, geo)


@pytest.mark.parametrize("N", [1, 3, 6])
def test_dof_counts(N):
    op = oracle_problem("poisson", N=N)
    assert op.ndof == (2 * N - 1) ** 2
    ns = oracle_problem("ns", N=N)
    assert ns.ndof == 2 * (2 * N - 1) ** 2 + (N + 1) ** 2 - 1


@pytest.mark.parametrize("N", [1, 2, 3, 7, 16, 37])
def test_product_mesh_equals_oracle_2d(N):
    for split, space in [("left", "C2"), ("left", "C1"), (False, "C2")]:
        a = om.rectangular_quad_mesh(N, split, space)
        b = pm.RectangularQuadMesh(N, split, space)
        assert a.kind == b.kind
        assert np.array_equal(a.coords, b.coords)       # bit for bit (N=37: inexact coords)
        assert np.array_equal(a.elements, b.elements)
        assert np.array_equal(a.on_boundary(), b.on_boundary())


@pytest.mark.parametrize("N", [1, 2, 4])
def test_product_mesh_equals_oracle_3d(N):
    a = om.cuboid_brick_mesh(N)
    b = pm.CuboidBrickMesh(N)
    assert np.array_equal(a.coords, b.coords)
    assert np.array_equal(a.elements, b.elements)
    assert np.array_equal(a.on_boundary(), b.on_boundary())


@pytest.mark.parametrize("kind,kw", [("poisson", dict(N=5)), ("ns", dict(N=4)),
                                     ("bratu", dict(N=7, space="C1", lam=1.0)),
                                     ("bratu", dict(N=3, space="C2", lam=1.0))])
def test_product_numbering_equals_oracle(kind, kw):
    mp = product_problem(kind, **kw)
    op = oracle_problem(kind, **kw)
    assert mp.ndof == op.ndof
    assert np.array_equal(mp.eqn, op.eqn)


def test_product_numbering_rd():
    mp = product_problem("rd", N=3)
    op = oracle_problem("rd", N=3, f=mp.equations.source_field)
    assert np.array_equal(mp.eqn, op.eqn)


def test_smoothed_noise_properties():
    mesh = pm.CuboidBrickMesh(3)
    f = wl.smoothed_noise(mesh, seed=4)
    assert f.shape == (mesh.nnode,)
    assert np.all(np.abs(f) <= 1.0)
    assert np.std(f) < np.std(np.random.default_rng(4).uniform(-1, 1, mesh.nnode))
