"""Pins the oracle's quadrature tables and shape functions against closed forms."""
import itertools
import math

import numpy as np
import pytest

from oracle import tables


def _tri_monomial(a, b):
    # integral of s0^a s1^b over the reference triangle {s0, s1 >= 0, s0 + s1 <= 1}
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)


@pytest.mark.parametrize("kind,degree,tol", [("tri_p1", 1, 2e-13), ("tri_p2", 5, 2e-12)])
def test_triangle_rules_integrate_polynomials(kind, degree, tol):
    K, W = tables.quadrature(kind)
    assert abs(W.sum() - 0.5) < tol
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            q = np.sum(W * K[:, 0] ** a * K[:, 1] ** b)
            assert abs(q - _tri_monomial(a, b)) < tol, (a, b)


def test_triangle_rule_is_not_exact_beyond_its_degree():
    K, W = tables.quadrature("tri_p2")
    q = np.sum(W * K[:, 0] ** 6)
    assert abs(q - _tri_monomial(6, 0)) > 1e-6


def test_gauss_2d_typo_is_kept():
    """oomph's Gauss<2,3> has 0.774596662941483 for the positive knot
    (integral.cc:84-93); sqrt(3/5) = 0.7745966692414834. The rule therefore integrates
    odd functions to a small non-zero value -- kept on purpose, it is the reference's
    table."""
    K, W = tables.quadrature("quad_q2")
    assert K.max() == 0.774596662941483
    assert K.min() == -0.774596669241483
    assert abs(W.sum() - 4.0) < 1e-15
    odd = np.sum(W * K[:, 0])
    assert 1e-9 < abs(odd) < 1e-8
    even = np.sum(W * K[:, 0] ** 2 * K[:, 1] ** 4)
    assert abs(even - (2 / 3) * (2 / 5)) < 5e-8


def test_gauss_3d_exact_to_degree_5():
    K, W = tables.quadrature("hex_q2")
    assert abs(W.sum() - 8.0) < 1e-12
    for a, b, c in itertools.product(range(6), repeat=3):
        if a > 5 or b > 5 or c > 5:
            continue
        exact = np.prod([0.0 if p % 2 else 2.0 / (p + 1) for p in (a, b, c)])
        q = np.sum(W * K[:, 0] ** a * K[:, 1] ** b * K[:, 2] ** c)
        assert abs(q - exact) < 1e-12, (a, b, c)


NODES = {
    "tri_p1": [(1, 0), (0, 1), (0, 0)],
    "tri_p2": [(1, 0), (0, 1), (0, 0), (0.5, 0.5), (0, 0.5), (0.5, 0)],
    "quad_q2": [(x, y) for y in (-1, 0, 1) for x in (-1, 0, 1)],
    "hex_q2": [(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)],
}


@pytest.mark.parametrize("kind", list(NODES))
def test_shape_kronecker_and_partition_of_unity(kind):
    nodes = NODES[kind]
    for i, s in enumerate(nodes):
        psi, _ = tables.shape(kind, s)
        e = np.zeros(len(nodes))
        e[i] = 1.0
        assert np.allclose(psi, e, atol=1e-15)
    rng = np.random.default_rng(0)
    for _ in range(20):
        s = rng.uniform(0, 0.5, size=len(nodes[0])) if kind.startswith("tri") else rng.uniform(-1, 1, size=len(nodes[0]))
        psi, d = tables.shape(kind, s)
        assert abs(psi.sum() - 1) < 1e-14
        assert np.allclose(d.sum(axis=0), 0, atol=1e-14)


@pytest.mark.parametrize("kind", list(NODES))
def test_shape_derivatives_match_finite_differences(kind):
    rng = np.random.default_rng(1)
    dim = len(NODES[kind][0])
    for _ in range(5):
        s = rng.uniform(0.1, 0.4, size=dim) if kind.startswith("tri") else rng.uniform(-0.9, 0.9, size=dim)
        _, d = tables.shape(kind, s)
        for k in range(dim):
            h = np.zeros(dim)
            h[k] = 1e-6
            fd = (tables.shape(kind, s + h)[0] - tables.shape(kind, s - h)[0]) / 2e-6
            assert np.allclose(d[:, k], fd, atol=1e-8)


def test_p2_reproduces_quadratics():
    nodes = NODES["tri_p2"]
    f = lambda s: 1 + 2 * s[0] - 3 * s[1] + s[0] * s[1] - 2 * s[0] ** 2 + 0.5 * s[1] ** 2
    vals = np.array([f(n) for n in nodes])
    for s in [(0.2, 0.3), (0.1, 0.7), (0.6, 0.05)]:
        psi, _ = tables.shape("tri_p2", s)
        assert abs(psi @ vals - f(s)) < 1e-14


def test_q1_pressure_basis_on_corners():
    for i, s in enumerate([(-1, -1), (1, -1), (-1, 1), (1, 1)]):
        psi, _ = tables.shape_q1_on_q2(s)
        e = np.zeros(4)
        e[i] = 1
        assert np.allclose(psi, e)
