"""Oracle restatement of pyoomph's structured mesh templates and oomph's numbering.

Test infrastructure only (see ``oracle/__init__``). Pure-Python loops: meant for meshes
the oracle finishes in seconds.

What is restated
----------------
* ``MeshTemplate.add_node_unique`` (src/meshtemplate.cpp:1378-1398): a node is created
  the first time a position is requested; later requests of the same position return the
  existing index. The reference finds equal positions with a kd-tree; here positions are
  matched on a 1e-9 grid, which is equivalent for the structured meshes used here.
* ``add_intermediate_node_unique`` with 2 / 4 parents (src/meshtemplate.cpp:1400-1418,
  1455-1490): position ``0.5*(a+b)`` resp. ``0.25*(a+b+c+d)``; the boundary set of a new
  node is the intersection of its parents' boundary sets (only when ``boundary_possible``
  for the 4-parent form).
* ``RectangularQuadMesh.define_geometry`` (pyoomph/meshes/simplemeshes.py:225-300) for
  ``split_in_tris`` in {False, "left"}, and ``CuboidBrickMesh.define_geometry``
  (simplemeshes.py:502-545).
* C1 -> C2 template conversion, node creation order included (src/meshtemplate.cpp:
  401-408 quads, 494-500 triangles, 579-621 bricks), performed element by element in
  element order when the mesh is built (src/meshtemplate.cpp:1188-1200).
* The oomph mesh keeps template node order (src/mesh2d.hpp:279-286, mesh3d.hpp:300-306),
  and ``Mesh::assign_global_eqn_numbers`` numbers nodes in that order, values in value
  order, skipping pinned values (oomph mesh.cc:686-708, nodes.cc:896-927).

Parity: pinned by tests/test_oracle_mesh.py (hand-counted node numbers of tiny meshes,
closed-form node / element / dof counts, Euler-characteristic and boundary counts).
"""
import numpy as np


class _Template:
    """Template node store with unique-position insertion."""

    def __init__(self, dim):
        self.dim = dim
        self.pos = []          # list of tuples (floats)
        self.bounds = []       # list of sets of boundary names
        self._lookup = {}

    @staticmethod
    def _key(p):
        return tuple(int(round(c * 1e9)) for c in p)

    def add_node_unique(self, p):
        k = self._key(p)
        idx = self._lookup.get(k)
        if idx is not None:
            return idx
        idx = len(self.pos)
        self._lookup[k] = idx
        self.pos.append(tuple(float(c) for c in p))
        self.bounds.append(set())
        return idx

    def add_intermediate2(self, n1, n2):
        a, b = self.pos[n1], self.pos[n2]
        p = tuple(0.5 * (a[i] + b[i]) for i in range(self.dim))
        ni = self.add_node_unique(p)
        if not self.bounds[ni]:
            self.bounds[ni] = set(self.bounds[n1]) & set(self.bounds[n2])
        return ni

    def add_intermediate4(self, n1, n2, n3, n4, boundary_possible):
        a, b, c, d = self.pos[n1], self.pos[n2], self.pos[n3], self.pos[n4]
        p = tuple(0.25 * (a[i] + b[i] + c[i] + d[i]) for i in range(self.dim))
        ni = self.add_node_unique(p)
        if boundary_possible and not self.bounds[ni]:
            self.bounds[ni] = (set(self.bounds[n1]) & set(self.bounds[n2])
                               & set(self.bounds[n3]) & set(self.bounds[n4]))
        return ni

    def add_to_boundary(self, name, nodes):
        for n in nodes:
            self.bounds[n].add(name)


class OracleMesh:
    """Nodes (template order), elements (oomph local node order) and boundary sets."""

    def __init__(self, kind, coords, elements, bounds):
        self.kind = kind                      # "tri_p1", "tri_p2", "quad_q2", "hex_q2"
        self.coords = np.asarray(coords, dtype=np.float64)
        self.elements = np.asarray(elements, dtype=np.int64)
        self.bounds = bounds

    @property
    def nnode(self):
        return self.coords.shape[0]

    @property
    def nelement(self):
        return self.elements.shape[0]

    def on_boundary(self, names=None):
        if names is None:
            return np.array([len(b) > 0 for b in self.bounds])
        names = set(names)
        return np.array([len(b & names) > 0 for b in self.bounds])


def rectangular_quad_mesh(N, split_in_tris=False, space="C2", size=1.0, lower_left=0.0):
    """Restates RectangularQuadMesh(N, size, lower_left, split_in_tris) followed by the
    C1/C2 element conversion (simplemeshes.py:225-300; meshtemplate.cpp:401-408,494-500).

    ``space`` is the dominant field space of the equations on the mesh: "C1" keeps the
    linear template elements, "C2" converts every element to its quadratic version.
    """
    Nx = Ny = int(N)
    sx = sy = float(size)
    llx = lly = float(lower_left)
    T = _Template(2)
    tmpl_elems = []  # ("tri", (a,b,c)) or ("quad", (n00,n10,n01,n11))
    for ix in range(Nx):
        for iy in range(Ny):
            n00 = T.add_node_unique((ix * sx / Nx + llx, iy * sy / Ny + lly))
            n10 = T.add_node_unique(((ix + 1) * sx / Nx + llx, iy * sy / Ny + lly))
            n01 = T.add_node_unique((ix * sx / Nx + llx, (iy + 1) * sy / Ny + lly))
            n11 = T.add_node_unique(((ix + 1) * sx / Nx + llx, (iy + 1) * sy / Ny + lly))
            if not split_in_tris:
                tmpl_elems.append(("quad", (n00, n10, n01, n11)))
            elif split_in_tris == "left":
                tmpl_elems.append(("tri", (n00, n10, n11)))
                tmpl_elems.append(("tri", (n00, n11, n01)))
            else:
                raise ValueError("only split_in_tris in {False,'left'} is restated")
            if ix == 0:
                T.add_to_boundary("left", [n00, n01])
            if ix == Nx - 1:
                T.add_to_boundary("right", [n10, n11])
            if iy == 0:
                T.add_to_boundary("bottom", [n00, n10])
            if iy == Ny - 1:
                T.add_to_boundary("top", [n01, n11])

    elements = []
    kind = None
    for typ, nn in tmpl_elems:
        if typ == "tri":
            if space == "C1":
                kind = "tri_p1"
                elements.append(list(nn))
            else:
                kind = "tri_p2"
                n3 = T.add_intermediate2(nn[0], nn[1])
                n4 = T.add_intermediate2(nn[1], nn[2])
                n5 = T.add_intermediate2(nn[0], nn[2])
                elements.append([nn[0], nn[1], nn[2], n3, n4, n5])
        else:
            if space == "C1":
                raise ValueError("Q1 quads are not on the hot path")
            kind = "quad_q2"
            n1 = T.add_intermediate2(nn[0], nn[1])
            n3 = T.add_intermediate2(nn[0], nn[2])
            n4 = T.add_intermediate4(nn[0], nn[1], nn[2], nn[3], False)
            n5 = T.add_intermediate2(nn[1], nn[3])
            n7 = T.add_intermediate2(nn[2], nn[3])
            elements.append([nn[0], n1, nn[1], n3, n4, n5, nn[2], n7, nn[3]])
    return OracleMesh(kind, T.pos, elements, T.bounds)


def cuboid_brick_mesh(N, size=1.0, lower_left=0.0):
    """Restates CuboidBrickMesh(N) with C2 conversion (simplemeshes.py:502-545;
    meshtemplate.cpp:579-621)."""
    N = int(N)
    s = float(size)
    ll = float(lower_left)
    T = _Template(3)
    bricks = []
    for ix in range(N):
        for iy in range(N):
            for iz in range(N):
                def add(px, py, pz):
                    return T.add_node_unique(((ix + px) * s / N + ll, (iy + py) * s / N + ll,
                                              (iz + pz) * s / N + ll))
                n000 = add(0, 0, 0); n100 = add(1, 0, 0); n010 = add(0, 1, 0); n110 = add(1, 1, 0)
                n001 = add(0, 0, 1); n101 = add(1, 0, 1); n011 = add(0, 1, 1); n111 = add(1, 1, 1)
                bricks.append((n000, n100, n010, n110, n001, n101, n011, n111))
                if ix == 0:
                    T.add_to_boundary("left", [n000, n010, n001, n011])
                if ix == N - 1:
                    T.add_to_boundary("right", [n100, n110, n101, n111])
                if iy == 0:
                    T.add_to_boundary("bottom", [n000, n100, n001, n101])
                if iy == N - 1:
                    T.add_to_boundary("top", [n010, n110, n011, n111])
                if iz == 0:
                    T.add_to_boundary("back", [n000, n100, n010, n110])
                if iz == N - 1:
                    T.add_to_boundary("front", [n001, n101, n011, n111])
    elements = []
    for v in bricks:
        n = [None] * 27
        n[0] = v[0]
        n[1] = T.add_intermediate2(v[0], v[1])
        n[2] = v[1]
        n[3] = T.add_intermediate2(v[0], v[2])
        n[4] = T.add_intermediate4(v[0], v[1], v[2], v[3], True)
        n[5] = T.add_intermediate2(v[1], v[3])
        n[6] = v[2]
        n[7] = T.add_intermediate2(v[2], v[3])
        n[8] = v[3]
        n[9] = T.add_intermediate2(v[0], v[4])
        n[11] = T.add_intermediate2(v[1], v[5])
        n[15] = T.add_intermediate2(v[2], v[6])
        n[17] = T.add_intermediate2(v[3], v[7])
        n[18] = v[4]
        n[19] = T.add_intermediate2(v[4], v[5])
        n[20] = v[5]
        n[21] = T.add_intermediate2(v[4], v[6])
        n[22] = T.add_intermediate4(v[4], v[5], v[6], v[7], True)
        n[23] = T.add_intermediate2(v[5], v[7])
        n[24] = v[6]
        n[25] = T.add_intermediate2(v[6], v[7])
        n[26] = v[7]
        n[10] = T.add_intermediate2(n[9], n[11])
        n[12] = T.add_intermediate2(n[9], n[15])
        n[14] = T.add_intermediate2(n[11], n[17])
        n[16] = T.add_intermediate2(n[15], n[17])
        n[13] = T.add_intermediate4(v[0], v[3], v[5], v[6], False)
        elements.append(n)
    return OracleMesh("hex_q2", T.pos, elements, T.bounds)


def number_equations(nnode, nvalue, pinned):
    """oomph ``Mesh::assign_global_eqn_numbers`` (mesh.cc:686-708) with
    ``Data::assign_eqn_numbers`` (nodes.cc:896-927): loop nodes in mesh order, values in
    value order; every value that is not pinned gets the next equation number.

    ``pinned``: bool array (nnode, nvalue). Returns int array (nnode, nvalue), -1 = pinned.
    """
    eqn = np.full((nnode, nvalue), -1, dtype=np.int64)
    k = 0
    for n in range(nnode):
        for v in range(nvalue):
            if not pinned[n, v]:
                eqn[n, v] = k
                k += 1
    return eqn, k
