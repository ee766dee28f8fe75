"""Oracle quadrature tables and Lagrange shape functions of the reference elements.

Test infrastructure only (see ``oracle/__init__``).

The quadrature tables are the reference's data, kept value for value (they are a table
that has to be alike: the 13/14-digit decimals and the mistyped knot of ``Gauss<2,3>``
change every element integral in the last digits):

* ``TGauss<2,2>`` (3 points)  oomph integral.cc:344-349  -- P1 triangles
* ``TGauss<2,3>`` (7 points)  oomph integral.cc:355-369  -- P2 triangles
* ``Gauss<2,3>``  (9 points)  oomph integral.cc:84-102   -- Q2 quadrilaterals
  (note ``0.774596662941483`` in the positive knots, as in the reference)
* ``Gauss<3,3>``  (27 points) oomph integral.cc:169-206  -- Q2 hexahedra

Default scheme per element: Telements.h:1515, Qelements.h:845,1257 (the pyoomph element
keeps it because ``integration_order`` stays 0: src/codegen.cpp:7059-7061,
pyoomph/generic/problem.py:759-763).

Shape functions (same arithmetic order as the reference):
* P1 triangle  oomph Telements.h:523-542
* P2 triangle  oomph Telements.h:627-662 (node order: s=(1,0),(0,1),(0,0),(.5,.5),(0,.5),(.5,0))
* 1D Lagrange  oomph shape.h:607-650 ; 2D/3D tensor products Qelements.cc:320-372, 621-658
* Q1 pressure sub-basis on a Q2 quad: src/elements.cpp:9288-9305 (C1 nodes 0,2,6,8)

Parity: pinned by tests/test_oracle_tables.py (partition of unity, Kronecker property at
nodes, derivative sums, exact integration of monomials up to the rule's order on the
exactly typed knots, weight sums, and the documented knot typo's effect).
"""
import numpy as np

_TG22_K = [[0.1666666666667, 0.1666666666667],
           [0.6666666666667, 0.1666666666667],
           [0.1666666666667, 0.6666666666667]]
_TG22_W = [0.1666666666667, 0.1666666666667, 0.1666666666667]

_TG23_K = [[0.1012865073235, 0.1012865073235],
           [0.7974269853531, 0.1012865073235],
           [0.1012865073235, 0.7974269853531],
           [0.4701420641051, 0.0597158717898],
           [0.4701420641051, 0.4701420641051],
           [0.0597158717898, 0.4701420641051],
           [0.3333333333333, 0.3333333333333]]
_TG23_W = [0.5 * 0.1259391805448, 0.5 * 0.1259391805448, 0.5 * 0.1259391805448,
           0.5 * 0.1323941527885, 0.5 * 0.1323941527885, 0.5 * 0.1323941527885,
           0.5 * 0.225]

_G23_K = [[-0.774596669241483, -0.774596669241483],
          [-0.774596669241483, 0.0],
          [-0.774596669241483, 0.774596662941483],
          [0.0, -0.774596669241483],
          [0.0, 0.0],
          [0.0, 0.774596662941483],
          [0.774596662941483, -0.774596669241483],
          [0.774596662941483, 0.0],
          [0.774596662941483, 0.774596662941483]]
_G23_W = [(25.0 / 81.0), (40.0 / 81.0), (25.0 / 81.0), (40.0 / 81.0), (64.0 / 81.0),
          (40.0 / 81.0), (25.0 / 81.0), (40.0 / 81.0), (25.0 / 81.0)]

_G33_a = 0.77459666924148
_G33_K = [[x, y, z] for x in (-_G33_a, 0.0, _G33_a) for y in (-_G33_a, 0.0, _G33_a)
          for z in (-_G33_a, 0.0, _G33_a)]
_G33_W = [0.17146776406035, 0.27434842249657, 0.17146776406035, 0.27434842249657,
          0.43895747599451, 0.27434842249657, 0.17146776406035, 0.27434842249657,
          0.17146776406035, 0.27434842249657, 0.43895747599451, 0.27434842249657,
          0.43895747599451, 0.70233196159122, 0.43895747599451, 0.27434842249657,
          0.43895747599451, 0.27434842249657, 0.17146776406035, 0.27434842249657,
          0.17146776406035, 0.27434842249657, 0.43895747599451, 0.27434842249657,
          0.17146776406035, 0.27434842249657, 0.17146776406035]


def quadrature(kind):
    """(knots [nqp, dim], weights [nqp]) of the default scheme of element ``kind``."""
    if kind == "tri_p1":
        return np.array(_TG22_K), np.array(_TG22_W)
    if kind == "tri_p2":
        return np.array(_TG23_K), np.array(_TG23_W)
    if kind == "quad_q2":
        return np.array(_G23_K), np.array(_G23_W)
    if kind == "hex_q2":
        return np.array(_G33_K), np.array(_G33_W)
    raise KeyError(kind)


def _lag3(s):
    return [0.5 * s * (s - 1.0), 1.0 - s * s, 0.5 * s * (s + 1.0)]


def _dlag3(s):
    return [s - 0.5, -2.0 * s, s + 0.5]


def shape(kind, s):
    """(psi [nnode], dpsids [nnode, dim]) at local coordinate ``s``."""
    s = [float(v) for v in s]
    if kind == "tri_p1":
        psi = [s[0], s[1], 1.0 - s[0] - s[1]]
        d = [[1.0, 0.0], [0.0, 1.0], [-1.0, -1.0]]
        return np.array(psi), np.array(d)
    if kind == "tri_p2":
        s2 = 1.0 - s[0] - s[1]
        psi = [2.0 * s[0] * (s[0] - 0.5), 2.0 * s[1] * (s[1] - 0.5), 2.0 * s2 * (s2 - 0.5),
               4.0 * s[0] * s[1], 4.0 * s[1] * s2, 4.0 * s2 * s[0]]
        c = 2.0 * (2.0 * s[0] - 1.5 + 2.0 * s[1])
        d = [[4.0 * s[0] - 1.0, 0.0], [0.0, 4.0 * s[1] - 1.0], [c, c],
             [4.0 * s[1], 4.0 * s[0]], [-4.0 * s[1], 4.0 * (1.0 - s[0] - 2.0 * s[1])],
             [4.0 * (1.0 - 2.0 * s[0] - s[1]), -4.0 * s[0]]]
        return np.array(psi), np.array(d)
    if kind == "quad_q2":
        P0, P1 = _lag3(s[0]), _lag3(s[1])
        D0, D1 = _dlag3(s[0]), _dlag3(s[1])
        psi, d = [], []
        for i in range(3):
            for j in range(3):
                psi.append(P1[i] * P0[j])
                d.append([P1[i] * D0[j], D1[i] * P0[j]])
        return np.array(psi), np.array(d)
    if kind == "hex_q2":
        P0, P1, P2 = _lag3(s[0]), _lag3(s[1]), _lag3(s[2])
        D0, D1, D2 = _dlag3(s[0]), _dlag3(s[1]), _dlag3(s[2])
        psi, d = [], []
        for i in range(3):
            for j in range(3):
                for k in range(3):
                    psi.append(P2[i] * P1[j] * P0[k])
                    d.append([P2[i] * P1[j] * D0[k], P2[i] * D1[j] * P0[k],
                              D2[i] * P1[j] * P0[k]])
        return np.array(psi), np.array(d)
    raise KeyError(kind)


def shape_q1_on_q2(s):
    """Bilinear pressure basis on the corner nodes 0,2,6,8 of a Q2 quad
    (src/elements.cpp:9288-9305; OneDimLagrange::shape<2>, shape.h:607-620)."""
    p1 = [0.5 * (1.0 - s[0]), 0.5 * (1.0 + s[0])]
    p2 = [0.5 * (1.0 - s[1]), 0.5 * (1.0 + s[1])]
    d1 = [-0.5, 0.5]
    d2 = [-0.5, 0.5]
    psi, d = [], []
    for i in range(2):
        for j in range(2):
            psi.append(p2[i] * p1[j])
            d.append([p2[i] * d1[j], d2[i] * p1[j]])
    return np.array(psi), np.array(d)


def tables(kind):
    """Reference tables at the quadrature points of the default scheme.

    Returns dict with w [nqp], psi [nqp, nnode], dpsi [nqp, nnode, dim] and, for the
    Q2 quad, psi1 [nqp, 4] (pressure sub-basis).
    """
    K, W = quadrature(kind)
    psi, dpsi, psi1 = [], [], []
    for q in range(len(W)):
        p, d = shape(kind, K[q])
        psi.append(p)
        dpsi.append(d)
        if kind == "quad_q2":
            psi1.append(shape_q1_on_q2(K[q])[0])
    out = {"w": np.array(W, dtype=np.float64), "knots": K,
           "psi": np.array(psi, dtype=np.float64), "dpsi": np.array(dpsi, dtype=np.float64)}
    if kind == "quad_q2":
        out["psi1"] = np.array(psi1, dtype=np.float64)
    return out
