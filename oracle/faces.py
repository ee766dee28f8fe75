"""Boundary-face (interface) elements of the scalar forms: the flux of pyoomph's
``NeumannBC(u=flux) @ boundary`` with flux = alpha*u + beta.

Test infrastructure only (see ``oracle/__init__``).

Restates, for P1 / P2 triangle, Q2 quadrilateral and Q2 brick meshes:

* which faces exist: ``Mesh`` interface elements are built on the bulk elements' faces
  that lie on the named boundary (src/mesh.cpp:615-660; ``InterfaceTElementLine1dC1`` /
  ``C2`` for ``BulkElementTri2dC1`` / ``C2``), one boundary after the other, bulk
  elements in element order, faces of an element in edge order (triangles 0-1, 1-2, 2-0;
  Q2 quads 0-1-2, 2-5-8, 6-7-8, 0-3-6; Q2 bricks k=0, k=2, j=0, j=2, i=0, i=2) --
  ``InterfaceElementLine1dC2`` for ``BulkElementQuad2dC2``, ``InterfaceElementQuad2dC2``
  for ``BulkElementBrick3dC2`` (``face_outputs_quad9``);
* the face element: a ``TElement<1,2>`` / ``TElement<1,3>`` line (nodes at s = 0, 1 or
  0, 1/2, 1; oomph Telements.h ``TElementShape<1,2>``, ``<1,3>`` shape and dshape_local)
  with its default ``TGauss<1,2>`` / ``TGauss<1,3>`` rule (oomph integral.cc:303-316; Q2
  quad faces: ``QElement<1,3>``, OneDimLagrange shapes on [-1, 1], ``Gauss<1,3>``;
  the generated functable's integration order is 0, so the element default is kept,
  src/mesh.cpp:664-667);
* the residual ``weak(flux, testfunction(u))`` (pyoomph/meshes/bcs.py:84-89): per point
  W = w * |dx/ds|, R_i += W * (flux * psi_i), J_ij += W * ((alpha * psi_j) * psi_i).

Pinned by closed forms in tests/test_oracle_faces.py: the edge integrals of 1 and of
linear data (P1: L/2 per node; P2: L * (1/6, 2/3, 1/6)), the 1D mass matrices
(L/6 [[2,1],[1,2]] and L/30 [[4,2,-1],[2,16,2],[-1,2,4]]), the Robin Jacobian as the
derivative of the residual (finite differences), and a Poisson problem whose exact
solution with a Neumann side is known.
"""
import numpy as np

# TGauss<1,2> / TGauss<1,3>: knots on [0, 1] and weights, as the reference writes them
_RULES = {
    2: (np.array([0.5 * (-0.577350269189626 + 1.0), 0.5 * (0.577350269189626 + 1.0)]), np.array([0.5, 0.5])),
    3: (np.array([0.5 * (-0.774596669241483 + 1.0), 0.5, 0.5 * (0.774596669241483 + 1.0)]),
        np.array([5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0])),
}

# Gauss<1,3> on [-1, 1] (integral.cc:50-53), the rule of QElement<1,3> faces of Q2 quads
_RULE_Q3 = (np.array([-0.774596669241483, 0.0, 0.774596669241483]), np.array([5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0]))

# element edges in local nodes, ends and (P2, Q2) midside node in face order
_EDGES = {"tri_p1": [(0, 1), (1, 2), (2, 0)], "tri_p2": [(0, 3, 1), (1, 4, 2), (2, 5, 0)],
          "quad_q2": [(0, 1, 2), (2, 5, 8), (6, 7, 8), (0, 3, 6)],
          # Q2 brick: local node i + 3j + 9k; a face's nodes a + 3b over its two free axes
          "hex_q2": [tuple(i + 3 * j for j in range(3) for i in range(3)),
                     tuple(18 + i + 3 * j for j in range(3) for i in range(3)),
                     tuple(i + 9 * k for k in range(3) for i in range(3)),
                     tuple(6 + i + 9 * k for k in range(3) for i in range(3)),
                     tuple(3 * j + 9 * k for k in range(3) for j in range(3)),
                     tuple(2 + 3 * j + 9 * k for k in range(3) for j in range(3))]}


def shape(nn, s, quad=False):
    """TElementShape<1,nn> (quad: OneDimLagrange shape<3>, dshape<3> of QElement<1,3>,
    shape.h:634-649): psi, dpsi/ds at local coordinate s."""
    if quad:
        return (np.array([0.5 * s * (s - 1.0), 1.0 - s * s, 0.5 * s * (s + 1.0)]),
                np.array([s - 0.5, -2.0 * s, s + 0.5]))
    if nn == 2:
        return np.array([1.0 - s, s]), np.array([-1.0, 1.0])
    psi = np.array([2.0 * (s - 1.0) * (s - 0.5), 4.0 * (1.0 - s) * s, 2.0 * (s - 0.5) * s])
    dpsi = np.array([4.0 * s - 3.0, 4.0 - 8.0 * s, 4.0 * s - 1.0])
    return psi, dpsi


def boundary_faces(mesh, names):
    """[nface, 2|3] node ids of the faces on each named boundary (in name order; bulk
    elements in order, edges 0-1, 1-2, 2-0)."""
    if mesh.kind not in _EDGES:
        raise ValueError("boundary faces are built for triangle, Q2 quadrilateral and Q2 brick meshes")
    faces = []
    for name in names:
        on = mesh.on_boundary([name])
        for e in range(mesh.nelement):
            el = mesh.elements[e]
            for edge in _EDGES[mesh.kind]:
                nodes = [int(el[k]) for k in edge]
                if all(on[n] for n in nodes):
                    faces.append(nodes)
    nn = len(_EDGES[mesh.kind][0])
    return np.array(faces, dtype=np.int64).reshape(-1, nn)


def face_outputs_quad9(X, U, alpha, beta, flag):
    """Faces of Q2 bricks (``InterfaceElementQuad2dC2``: ``QElement<2,3>`` with the Q2
    quadrilateral's ``Gauss<2,3>`` table, its mistyped knot included): X [nf, 9, 3], U [nf, 9].
    Per point W = w |t_0 x t_1| and the flux come first; each entry then sums its nine point
    terms in point order."""
    from .tables import tables
    T = tables("quad_q2")
    w, psi, dpsi = T["w"], T["psi"], T["dpsi"]
    nf = U.shape[0]
    nq = len(w)
    W = np.zeros((nf, nq))
    Fl = np.zeros((nf, nq))
    for q in range(nq):
        t0 = np.zeros((nf, 3))
        t1 = np.zeros((nf, 3))
        u = np.zeros(nf)
        for l in range(9):
            for d in range(3):
                t0[:, d] = t0[:, d] + X[:, l, d] * dpsi[q, l, 0]
                t1[:, d] = t1[:, d] + X[:, l, d] * dpsi[q, l, 1]
            u = u + U[:, l] * psi[q, l]
        n0 = t0[:, 1] * t1[:, 2] - t0[:, 2] * t1[:, 1]
        n1 = t0[:, 2] * t1[:, 0] - t0[:, 0] * t1[:, 2]
        n2 = t0[:, 0] * t1[:, 1] - t0[:, 1] * t1[:, 0]
        W[:, q] = w[q] * np.sqrt((n0 * n0 + n1 * n1) + n2 * n2)
        Fl[:, q] = alpha * u + beta
    R = np.zeros((nf, 9))
    J = np.zeros((nf, 9, 9))
    for i in range(9):
        r = np.zeros(nf)
        for q in range(nq):
            r = r + W[:, q] * (Fl[:, q] * psi[q, i])
        R[:, i] = r
        if flag == 1:
            for j in range(9):
                v = np.zeros(nf)
                for q in range(nq):
                    v = v + W[:, q] * ((alpha * psi[q, j]) * psi[q, i])
                J[:, i, j] = v
    return R, J


def face_outputs(X, U, alpha, beta, flag, quad=False):
    """Element residuals [nf, nn] and Jacobians [nf, nn, nn] in face-node numbering.

    X [nf, nn, 2] node coordinates, U [nf, nn] nodal values, alpha / beta scalars or one
    value per face; quad: the faces of Q2 quadrilaterals."""
    nf, nn = U.shape
    knots, weights = _RULE_Q3 if quad else _RULES[nn]
    R = np.zeros((nf, nn))
    J = np.zeros((nf, nn, nn))
    for g in range(nn):
        psi, dpsi = shape(nn, knots[g], quad)
        tx = np.zeros(nf)
        ty = np.zeros(nf)
        u = np.zeros(nf)
        for l in range(nn):
            tx = tx + X[:, l, 0] * dpsi[l]
            ty = ty + X[:, l, 1] * dpsi[l]
            u = u + U[:, l] * psi[l]
        W = weights[g] * np.sqrt(tx * tx + ty * ty)
        flux = alpha * u + beta
        for i in range(nn):
            R[:, i] = R[:, i] + W * (flux * psi[i])
            if flag == 1:
                for j in range(nn):
                    J[:, i, j] = J[:, i, j] + W * ((alpha * psi[j]) * psi[i])
    return R, J


def enforced_face_outputs(X, U, Lm, alpha, beta, flag, quad=False, cexpr=None, C=None, deriv=False):
    """Faces of pyoomph's EnforcedBC(u=alpha*u + beta) @ boundary (pyoomph/meshes/bcs.py:
    129-253): the Lagrange multiplier l (value 1 of the face nodes, u's space) adds the
    residuals weak(alpha*u + beta, l_test) and weak(l, u_test) (bcs.py:185-189). Element
    residuals [nf, 2nn] and Jacobians [nf, 2nn, 2nn] in the slot order (u_0, l_0, u_1, l_1,
    ...) of the face nodes' values; X [nf, nn, 2], U / Lm [nf, nn] nodal u and l, alpha /
    beta scalars or one value per face; the rules and geometry of face_outputs. Per point:
    R[u_i] += W (l psi_i), R[l_i] += W ((alpha u + beta) psi_i), J[u_i][l_j] += W (psi_j
    psi_i), J[l_i][u_j] += W ((alpha psi_j) psi_i); the u-u and l-l blocks stay zero.
    With cexpr = (c, dc/du, dc/dp, d2c/du dp) (functions of u, x0, x1, C; the last two may
    be None) the constraint is the expression c with C [nf, ncoef]; with deriv, the
    parameter derivatives: the u rows vanish, the l rows take dc/dp and d2c/du dp."""
    nf, nn = U.shape
    knots, weights = _RULE_Q3 if quad else _RULES[nn]
    R = np.zeros((nf, 2 * nn))
    J = np.zeros((nf, 2 * nn, 2 * nn))
    for g in range(nn):
        psi, dpsi = shape(nn, knots[g], quad)
        tx = np.zeros(nf)
        ty = np.zeros(nf)
        u = np.zeros(nf)
        lm = np.zeros(nf)
        for l in range(nn):
            tx = tx + X[:, l, 0] * dpsi[l]
            ty = ty + X[:, l, 1] * dpsi[l]
            u = u + U[:, l] * psi[l]
            lm = lm + Lm[:, l] * psi[l]
        W = weights[g] * np.sqrt(tx * tx + ty * ty)
        if cexpr is None:
            c, dc = alpha * u + beta, alpha
        else:
            x0 = np.zeros(nf)
            x1 = np.zeros(nf)
            for l in range(nn):
                x0 = x0 + X[:, l, 0] * psi[l]
                x1 = x1 + X[:, l, 1] * psi[l]
            fc, fd = (cexpr[2], cexpr[3]) if deriv else (cexpr[0], cexpr[1])
            c = fc(u, x0, x1, C) if fc is not None else np.zeros(nf)
            dc = fd(u, x0, x1, C) if fd is not None else np.zeros(nf)
        for i in range(nn):
            R[:, 2 * i] = R[:, 2 * i] + W * (lm * psi[i])
            R[:, 2 * i + 1] = R[:, 2 * i + 1] + W * (c * psi[i])
            if flag == 1:
                for j in range(nn):
                    J[:, 2 * i, 2 * j + 1] = J[:, 2 * i, 2 * j + 1] + W * (psi[j] * psi[i])
                    J[:, 2 * i + 1, 2 * j] = J[:, 2 * i + 1, 2 * j] + W * ((dc * psi[j]) * psi[i])
    if deriv:
        R[:, 0::2] = 0.0
        J[:, 0::2, :] = 0.0
    return R, J


def expression_face_outputs(X, U, C, flux, dflux, flag, quad=False, sides=None):
    """Faces of NeumannBC(u=flux) with a flux expression (the generated boundary-face forms):
    flux(u, x0, x1, C) and its u-derivative dflux(...) evaluated per point for all faces at
    once (C [nf, ncoef] the faces' coefficients); otherwise face_outputs' rules and order:
    R_i += W (flux psi_i), J_ij += W ((dflux psi_j) psi_i). With sides [nf] (+-1), the
    expressions also take the outward unit normal (n0, n1): the line element's normal of
    the reference, the tangent turned counter-clockwise over its length (src/elements.cpp:
    1724-1750), times the face's side."""
    nf, nn = U.shape
    knots, weights = _RULE_Q3 if quad else _RULES[nn]
    R = np.zeros((nf, nn))
    J = np.zeros((nf, nn, nn))
    for g in range(nn):
        psi, dpsi = shape(nn, knots[g], quad)
        tx = np.zeros(nf)
        ty = np.zeros(nf)
        u = np.zeros(nf)
        x0 = np.zeros(nf)
        x1 = np.zeros(nf)
        for l in range(nn):
            tx = tx + X[:, l, 0] * dpsi[l]
            ty = ty + X[:, l, 1] * dpsi[l]
            u = u + U[:, l] * psi[l]
        for l in range(nn):
            x0 = x0 + X[:, l, 0] * psi[l]
        for l in range(nn):
            x1 = x1 + X[:, l, 1] * psi[l]
        W = weights[g] * np.sqrt(tx * tx + ty * ty)
        if sides is None:
            F = flux(u, x0, x1, C)
            D = dflux(u, x0, x1, C)
        else:
            tl = np.sqrt(tx * tx + ty * ty)
            n0, n1 = sides * (-ty) / tl, sides * tx / tl
            F = flux(u, x0, x1, C, n0, n1)
            D = dflux(u, x0, x1, C, n0, n1)
        for i in range(nn):
            R[:, i] = R[:, i] + W * (F * psi[i])
            if flag == 1:
                for j in range(nn):
                    J[:, i, j] = J[:, i, j] + W * ((D * psi[j]) * psi[i])
    return R, J


def outward_sides(coords, elements, faces):
    """+-1 per line face so that the reference's line normal (tangent from the first to the
    last face node turned counter-clockwise) leaves the bulk element holding the face (found
    by its nodes; its node centroid is inside it)."""
    coords = np.asarray(coords)[:, :2]
    elements = np.asarray(elements)
    holder = {}
    for e, nodes in enumerate(elements):
        for n in nodes:
            holder.setdefault(int(n), set()).add(e)
    sides = np.zeros(len(faces))
    for f, fn in enumerate(faces):
        common = set.intersection(*[holder[int(n)] for n in fn])
        e = min(common)
        t = coords[fn[-1]] - coords[fn[0]]
        n = np.array([-t[1], t[0]])
        mid = 0.5 * (coords[fn[-1]] + coords[fn[0]])
        sides[f] = 1.0 if np.dot(n, mid - coords[elements[e]].mean(axis=0)) > 0.0 else -1.0
    return sides
