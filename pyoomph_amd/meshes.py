"""Structured meshes with pyoomph's node and element numbering.

Mirrors ``RectangularQuadMesh`` and ``CuboidBrickMesh`` (pyoomph/meshes/simplemeshes.py:
225-300, 456-545) together with the C1->C2 template conversion and the node order the
oomph mesh inherits from the template (src/meshtemplate.cpp:401-408, 494-500, 579-621,
1188-1200, 1378-1418; src/mesh2d.hpp:265-300, mesh3d.hpp:285-320).

The reference inserts nodes one at a time (``add_node_unique`` finds an existing node at
the same position). Here the same sequence of node requests is generated as arrays, in
the reference's request order, and every request is mapped to the first request of its
position (``numpy.unique`` on integer keys of a doubled lattice): the node number of a
position is the rank of its first request -- the same numbering, without a per-node
search. A node keeps the coordinates computed by its first request (vertices
``i*size/N + lower_left``, edge midpoints ``0.5*(a+b)``, face / cell centres
``0.25*(a+b+c+d)`` or midpoints of edge midpoints), as in the reference.
"""
import numpy as np

__all__ = ["StructuredMesh", "RectangularQuadMesh", "CuboidBrickMesh", "boundary_faces"]


class StructuredMesh:
    """Built mesh: node coordinates, element connectivity and boundary membership.

    kind        "tri_p1", "tri_p2", "quad_q2" or "hex_q2"
    coords      float64 [nnode, dim]
    elements    int32 [nelement, nnode_per_element] (oomph local node order)
    lattice     int64 [nnode, dim] doubled-lattice position (0 .. 2N per axis)
    """

    BOUNDARY_AXES = {"left": (0, 0), "right": (0, 1), "bottom": (1, 0), "top": (1, 1),
                     "back": (2, 0), "front": (2, 1)}

    def __init__(self, kind, coords, elements, lattice, N):
        self.kind = kind
        self.coords = coords
        self.elements = elements
        self.lattice = lattice
        self.N = N

    @property
    def nnode(self):
        return self.coords.shape[0]

    @property
    def nelement(self):
        return self.elements.shape[0]

    @property
    def dim(self):
        return self.coords.shape[1]

    def boundary_nodes(self, name):
        """Bool mask of the nodes on boundary ``name``."""
        ax, side = self.BOUNDARY_AXES[name]
        if ax >= self.dim:
            return np.zeros(self.nnode, dtype=bool)
        return self.lattice[:, ax] == (0 if side == 0 else 2 * self.N)

    def on_boundary(self, names=None):
        if names is None:
            names = [n for n, (ax, _) in self.BOUNDARY_AXES.items() if ax < self.dim]
        m = np.zeros(self.nnode, dtype=bool)
        for n in names:
            m |= self.boundary_nodes(n)
        return m


# element edges in local nodes: ends and (P2, Q2) midside node, in face-node order
_FACE_EDGES = {"tri_p1": ((0, 1), (1, 2), (2, 0)), "tri_p2": ((0, 3, 1), (1, 4, 2), (2, 5, 0)),
               "quad_q2": ((0, 1, 2), (2, 5, 8), (6, 7, 8), (0, 3, 6)),
               # Q2 brick faces (local node i + 3j + 9k), face node a + 3b over the face's two axes
               "hex_q2": ((0, 1, 2, 3, 4, 5, 6, 7, 8), (18, 19, 20, 21, 22, 23, 24, 25, 26),
                          (0, 1, 2, 9, 10, 11, 18, 19, 20), (6, 7, 8, 15, 16, 17, 24, 25, 26),
                          (0, 3, 6, 9, 12, 15, 18, 21, 24), (2, 5, 8, 11, 14, 17, 20, 23, 26))}


def boundary_faces(mesh, names, return_elements=False):
    """Nodes of the bulk faces on the named boundaries, for the interface elements that
    carry a boundary flux (the reference builds one interface element per bulk face on
    the boundary, src/mesh.cpp:615-660): int32 [nface, 2] (P1 triangles: end, end) or
    [nface, 3] (P2 triangles, Q2 quadrilaterals: end, midside, end) or [nface, 9] (Q2 bricks:
    the face's 3x3 node lattice); one boundary after the other, bulk elements in element
    order, faces of each in the order of ``_FACE_EDGES``. With return_elements, also the bulk
    element of each face."""
    if mesh.kind not in _FACE_EDGES:
        raise ValueError("boundary-face fluxes are built for P1 / P2 triangle, Q2 quadrilateral and Q2 brick meshes")
    E = mesh.elements
    edges = _FACE_EDGES[mesh.kind]
    out, owners = [], []
    for name in names:
        on = mesh.boundary_nodes(name)
        sel = np.stack([np.all(on[E[:, list(ed)]], axis=1) for ed in edges], axis=1)   # [nel, 3]
        el, k = np.nonzero(sel)                                                        # element order, edge order
        f = np.stack([E[el, ed] for ed in np.array(edges)[k].T], axis=1) if len(el) else \
            np.zeros((0, len(edges[0])), dtype=np.int64)
        out.append(f)
        owners.append(el)
    faces = np.ascontiguousarray(np.concatenate(out), dtype=np.int32)
    c = mesh.coords
    if faces.shape[1] == 3 and len(faces):
        if not np.allclose(c[faces[:, 1]], 0.5 * (c[faces[:, 0]] + c[faces[:, 2]]), rtol=0, atol=1e-12):
            raise ValueError("midside node of a boundary face is not at the edge midpoint")
    if faces.shape[1] == 9 and len(faces):
        ok = np.allclose(c[faces[:, 1]], 0.5 * (c[faces[:, 0]] + c[faces[:, 2]]), rtol=0, atol=1e-12) and \
            np.allclose(c[faces[:, 4]], 0.25 * (c[faces[:, 0]] + c[faces[:, 2]] + c[faces[:, 6]] + c[faces[:, 8]]),
                        rtol=0, atol=1e-12)
        if not ok:
            raise ValueError("boundary face nodes of a Q2 brick are not in face-lattice order")
    if return_elements:
        return faces, np.concatenate(owners).astype(np.int64)
    return faces


def face_sides(mesh, faces, elements):
    """+1 or -1 per line face: the side whose clockwise-turned tangent (last node minus
    first) leaves its bulk element (points away from the element's node centroid)."""
    c = np.asarray(mesh.coords)[:, :2]
    t = c[faces[:, -1]] - c[faces[:, 0]]
    n = np.stack([t[:, 1], -t[:, 0]], axis=1)
    away = 0.5 * (c[faces[:, -1]] + c[faces[:, 0]]) - c[np.asarray(mesh.elements)[elements]].mean(axis=1)
    return np.where(np.sum(n * away, axis=1) > 0.0, 1.0, -1.0)


def _number(keys, coords):
    """Node number of every request (rank of the first request of its key)."""
    uniq, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty(len(uniq), dtype=np.int64)
    rank[order] = np.arange(len(uniq))
    node_of_request = rank[inv.reshape(-1)]
    node_coords = coords[first[order]]
    return node_of_request, node_coords, order, first


def RectangularQuadMesh(N, split_in_tris=False, space="C2", size=1.0, lower_left=0.0):
    """Unit-square style mesh of N x N cells (``split_in_tris`` False or "left").

    ``space`` is the dominant space of the equations: "C1" keeps linear triangles, "C2"
    converts every cell to its quadratic element.
    """
    N = int(N)
    if N <= 0:
        raise ValueError("Mesh size must be a positive integer")
    size, ll = float(size), float(lower_left)
    ix, iy = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")   # ix outer, iy inner
    ix, iy = ix.reshape(-1), iy.reshape(-1)
    # vertex requests per cell: n00, n10, n01, n11
    px = np.array([0, 1, 0, 1])
    py = np.array([0, 0, 1, 1])
    VX = (ix[:, None] + px[None, :])       # [ncell, 4] lattice vertex index
    VY = (iy[:, None] + py[None, :])
    vx = (VX * size) / N + ll
    vy = (VY * size) / N + ll
    req_key = [(2 * VX) * (2 * N + 1) + 2 * VY]
    req_xy = [np.stack([vx, vy], axis=-1)]
    ncell = N * N
    vid = np.arange(ncell * 4).reshape(ncell, 4)   # request index of each cell vertex
    if split_in_tris:
        if split_in_tris != "left":
            raise ValueError("only split_in_tris in {False, 'left'} is supported")
        tri = np.stack([vid[:, [0, 1, 3]], vid[:, [0, 3, 2]]], axis=1).reshape(-1, 3)  # request ids
        if space == "C1":
            keys = req_key[0].reshape(-1)
            xy = req_xy[0].reshape(-1, 2)
            node_of_req, node_xy, order, first = _number(keys, xy)
            lat = np.stack([(keys[first[order]] // (2 * N + 1)), keys[first[order]] % (2 * N + 1)], axis=1)
            return StructuredMesh("tri_p1", node_xy, node_of_req[tri].astype(np.int32), lat, N)
        keys0 = req_key[0].reshape(-1)
        xy0 = req_xy[0].reshape(-1, 2)
        a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
        mids = [(a, b), (b, c), (a, c)]                # n3, n4, n5 (meshtemplate.cpp:494-500)
        mk = np.stack([(keys0[u] + keys0[v]) // 2 for u, v in mids], axis=1)     # [ntri, 3]
        mxy = np.stack([0.5 * (xy0[u] + xy0[v]) for u, v in mids], axis=1)      # [ntri, 3, 2]
        keys = np.concatenate([keys0, mk.reshape(-1)])
        xy = np.concatenate([xy0, mxy.reshape(-1, 2)])
        node_of_req, node_xy, order, first = _number(keys, xy)
        nv = len(keys0)
        mid_ids = nv + np.arange(mk.size).reshape(-1, 3)
        conn = np.concatenate([node_of_req[tri], node_of_req[mid_ids]], axis=1)
        k = keys[first[order]]
        lat = np.stack([k // (2 * N + 1), k % (2 * N + 1)], axis=1)
        return StructuredMesh("tri_p2", node_xy, conn.astype(np.int32), lat, N)
    if space == "C1":
        raise ValueError("Q1 quadrilaterals are not supported")
    keys0 = req_key[0].reshape(-1)
    xy0 = req_xy[0].reshape(-1, 2)
    q = vid  # [ncell, 4] (n00, n10, n01, n11)
    # meshtemplate.cpp:401-408: n1=(0,1) n3=(0,2) n4=centre(0,1,2,3) n5=(1,3) n7=(2,3)
    def mid(u, v):
        return (keys0[u] + keys0[v]) // 2, 0.5 * (xy0[u] + xy0[v])
    k1, x1 = mid(q[:, 0], q[:, 1])
    k3, x3 = mid(q[:, 0], q[:, 2])
    k4 = (keys0[q[:, 0]] + keys0[q[:, 3]]) // 2
    x4 = 0.25 * (xy0[q[:, 0]] + xy0[q[:, 1]] + xy0[q[:, 2]] + xy0[q[:, 3]])
    k5, x5 = mid(q[:, 1], q[:, 3])
    k7, x7 = mid(q[:, 2], q[:, 3])
    mk = np.stack([k1, k3, k4, k5, k7], axis=1)
    mxy = np.stack([x1, x3, x4, x5, x7], axis=1)
    keys = np.concatenate([keys0, mk.reshape(-1)])
    xy = np.concatenate([xy0, mxy.reshape(-1, 2)])
    node_of_req, node_xy, order, first = _number(keys, xy)
    nv = len(keys0)
    m_ids = nv + np.arange(mk.size).reshape(-1, 5)
    V = node_of_req[q]
    M = node_of_req[m_ids]
    conn = np.stack([V[:, 0], M[:, 0], V[:, 1], M[:, 1], M[:, 2], M[:, 3], V[:, 2], M[:, 4],
                     V[:, 3]], axis=1)
    k = keys[first[order]]
    lat = np.stack([k // (2 * N + 1), k % (2 * N + 1)], axis=1)
    return StructuredMesh("quad_q2", node_xy, conn.astype(np.int32), lat, N)


def CuboidBrickMesh(N, size=1.0, lower_left=0.0):
    """N x N x N cube of Q2 bricks (CuboidBrickMesh + C2 conversion)."""
    N = int(N)
    if N <= 0:
        raise ValueError("Mesh size must be a positive integer")
    size, ll = float(size), float(lower_left)
    L = 2 * N + 1
    ix, iy, iz = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    ix, iy, iz = ix.reshape(-1), iy.reshape(-1), iz.reshape(-1)
    ncell = ix.size
    # n000 n100 n010 n110 n001 n101 n011 n111 (simplemeshes.py:519-528)
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1],
                  [1, 1, 1]])
    VX = ix[:, None] + p[None, :, 0]
    VY = iy[:, None] + p[None, :, 1]
    VZ = iz[:, None] + p[None, :, 2]
    keys0 = (((2 * VX) * L + 2 * VY) * L + 2 * VZ).reshape(-1)
    xyz0 = np.stack([(VX * size) / N + ll, (VY * size) / N + ll, (VZ * size) / N + ll],
                    axis=-1).reshape(-1, 3)
    v = np.arange(ncell * 8).reshape(ncell, 8)

    def mid2(ka, xa, kb, xb):
        return (ka + kb) // 2, 0.5 * (xa + xb)

    def vk(i):
        return keys0[v[:, i]], xyz0[v[:, i]]

    def mid_v(i, j):
        (ka, xa), (kb, xb) = vk(i), vk(j)
        return mid2(ka, xa, kb, xb)

    def centre4(i, j, k, l):
        ks = [keys0[v[:, t]] for t in (i, j, k, l)]
        xs = [xyz0[v[:, t]] for t in (i, j, k, l)]
        return (ks[0] + ks[1] + ks[2] + ks[3]) // 4, 0.25 * (xs[0] + xs[1] + xs[2] + xs[3])

    # request order of meshtemplate.cpp:579-621
    reqs = {}
    reqs[1] = mid_v(0, 1)
    reqs[3] = mid_v(0, 2)
    reqs[4] = centre4(0, 1, 2, 3)
    reqs[5] = mid_v(1, 3)
    reqs[7] = mid_v(2, 3)
    reqs[9] = mid_v(0, 4)
    reqs[11] = mid_v(1, 5)
    reqs[15] = mid_v(2, 6)
    reqs[17] = mid_v(3, 7)
    reqs[19] = mid_v(4, 5)
    reqs[21] = mid_v(4, 6)
    reqs[22] = centre4(4, 5, 6, 7)
    reqs[23] = mid_v(5, 7)
    reqs[25] = mid_v(6, 7)
    reqs[10] = mid2(*reqs[9], *reqs[11])
    reqs[12] = mid2(*reqs[9], *reqs[15])
    reqs[14] = mid2(*reqs[11], *reqs[17])
    reqs[16] = mid2(*reqs[15], *reqs[17])
    # body centre 0.25*(v0+v3+v5+v6): key = average of the 4 (all even lattice sums)
    k13 = (keys0[v[:, 0]] + keys0[v[:, 3]] + keys0[v[:, 5]] + keys0[v[:, 6]]) // 4
    x13 = 0.25 * (xyz0[v[:, 0]] + xyz0[v[:, 3]] + xyz0[v[:, 5]] + xyz0[v[:, 6]])
    reqs[13] = (k13, x13)
    order_req = [1, 3, 4, 5, 7, 9, 11, 15, 17, 19, 21, 22, 23, 25, 10, 12, 14, 16, 13]
    mk = np.stack([reqs[i][0] for i in order_req], axis=1)
    mx = np.stack([reqs[i][1] for i in order_req], axis=1)
    keys = np.concatenate([keys0, mk.reshape(-1)])
    xyz = np.concatenate([xyz0, mx.reshape(-1, 3)])
    node_of_req, node_xyz, order, first = _number(keys, xyz)
    nv = len(keys0)
    m_ids = nv + np.arange(mk.size).reshape(-1, len(order_req))
    V = node_of_req[v]
    M = node_of_req[m_ids]
    conn = np.empty((ncell, 27), dtype=np.int64)
    for slot, vi in zip([0, 2, 6, 8, 18, 20, 24, 26], range(8)):
        conn[:, slot] = V[:, vi]
    for col, slot in enumerate(order_req):
        conn[:, slot] = M[:, col]
    k = keys[first[order]]
    lat = np.stack([k // (L * L), (k // L) % L, k % L], axis=1)
    return StructuredMesh("hex_q2", node_xyz, conn.astype(np.int32), lat, N)
