"""The multigrid's element level (host code, no GPU): on P2 triangles, Q2 quadrilaterals and
Q2 bricks of a scalar form, the first coarse level is the space of the linear (P1 / Q1)
functions on the element vertices. Pinned by what the mathematics fixes: the interpolation
reproduces every linear (bilinear, trilinear) function at every node exactly, a pinned
vertex only removes its column, and the plan's patterns are scipy's sparse products of that
interpolation."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

from pyoomph_amd import _lib
from pyoomph_amd.meshes import CuboidBrickMesh, RectangularQuadMesh
from tests.test_amg_plan import _bool, _same_pattern


def _rows(mesh, eqn, ndof, form=_lib.FORM_POISSON, nvalue=1):
    L = _lib.load()
    L.pyhip_element_level_rows.argtypes = [ctypes.POINTER(_lib.SubproblemDesc), _lib.c_i64p, _lib.c_i32p,
                                           _lib.c_i32p, _lib.c_dp, _lib.c_i64p]
    conn = np.ascontiguousarray(mesh.elements, np.int32)
    eqn = np.ascontiguousarray(eqn, np.int32)
    d = _lib.SubproblemDesc()
    d.form, d.elem_kind = form, _lib.ELEM_KIND[mesh.kind]
    d.nnode, d.nelem, d.nvalue, d.ndof = mesh.nnode, mesh.nelement, nvalue, ndof
    d.conn = conn.ctypes.data_as(_lib.c_i32p)
    d.eqn = eqn.ctypes.data_as(_lib.c_i32p)
    nc, nnz = ctypes.c_int64(), ctypes.c_int64()
    rs = np.zeros(ndof + 1, np.int32)
    _lib.check(L.pyhip_element_level_rows(ctypes.byref(d), ctypes.byref(nc), rs.ctypes.data_as(_lib.c_i32p),
                                          None, None, ctypes.byref(nnz)))
    col = np.zeros(max(nnz.value, 1), np.int32)
    w = np.zeros(max(nnz.value, 1))
    _lib.check(L.pyhip_element_level_rows(ctypes.byref(d), ctypes.byref(nc), rs.ctypes.data_as(_lib.c_i32p),
                                          col.ctypes.data_as(_lib.c_i32p), w.ctypes.data_as(_lib.c_dp),
                                          ctypes.byref(nnz)))
    return nc.value, rs, col[:nnz.value], w[:nnz.value]


MESHES = [("tri_p2", lambda: RectangularQuadMesh(5, split_in_tris="left", space="C2")),
          ("quad_q2", lambda: RectangularQuadMesh(4, space="C2")),
          ("hex_q2", lambda: CuboidBrickMesh(3))]


def _vertices(mesh):
    nv = {"tri_p2": [0, 1, 2], "quad_q2": [0, 2, 6, 8], "hex_q2": [0, 2, 6, 8, 18, 20, 24, 26]}[mesh.kind]
    return np.unique(np.asarray(mesh.elements)[:, nv])


@pytest.mark.parametrize("kind,make", MESHES)
def test_interpolation_reproduces_multilinear_functions(kind, make):
    mesh = make()
    assert mesh.kind == kind
    n = mesh.nnode
    nc, rs, col, w = _rows(mesh, np.arange(n), n)
    verts = _vertices(mesh)
    assert nc == len(verts) and nc < n
    P = sp.csr_matrix((w, col, rs), shape=(n, nc))
    X = np.asarray(mesh.coords)[:, :mesh.coords.shape[1]]
    dim = X.shape[1]
    rng = np.random.default_rng(2)
    # linear on triangles; bilinear / trilinear (every product of distinct coordinates) on Q2
    terms = [np.ones(n)] + [X[:, a] for a in range(dim)]
    if kind != "tri_p2":
        terms += [X[:, a] * X[:, b] for a in range(dim) for b in range(a + 1, dim)]
        if dim == 3:
            terms.append(X[:, 0] * X[:, 1] * X[:, 2])
    for _ in range(3):
        f = sum(rng.uniform(-1, 1) * t for t in terms)
        assert np.max(np.abs(P @ f[verts] - f)) < 1e-14
    # a vertex row is the vertex itself
    assert np.array_equal(np.diff(rs)[verts], np.ones(nc, int)) and np.all(w[rs[verts]] == 1.0)
    # but not a quadratic function (the level is coarser)
    g = X[:, 0] ** 2
    assert np.max(np.abs(P @ g[verts] - g)) > 1e-3


def test_pinned_vertices_drop_their_columns():
    mesh = RectangularQuadMesh(5, split_in_tris="left", space="C2")
    n = mesh.nnode
    nc0, rs0, col0, w0 = _rows(mesh, np.arange(n), n)
    pinned = mesh.on_boundary({"left", "bottom"})
    eqn = np.full(n, -1)
    eqn[~pinned] = np.arange((~pinned).sum())
    nc, rs, col, w = _rows(mesh, eqn, int((~pinned).sum()))
    verts = _vertices(mesh)
    vfree = verts[~pinned[verts]]
    assert nc == len(vfree)
    full = sp.csr_matrix((w0, col0, rs0), shape=(n, nc0)).toarray()
    # columns: the free vertices in node order; rows: the unknowns
    want = full[~pinned][:, ~pinned[verts]]
    got = sp.csr_matrix((w, col, rs), shape=(int((~pinned).sum()), nc)).toarray()
    assert np.array_equal(got, want)


def test_linear_triangles_and_mixed_forms_have_no_element_level():
    mesh = RectangularQuadMesh(4, split_in_tris="left", space="C1")
    nc, rs, col, w = _rows(mesh, np.arange(mesh.nnode), mesh.nnode)
    assert nc == 0 and len(col) == 0


def _plan_fixed(A, off, fix, cap=16, bottom=20):
    L = _lib.load()
    L.pyhip_amg_plan_new_fixed.restype = ctypes.c_void_p
    L.pyhip_amg_plan_new_fixed.argtypes = [ctypes.c_int64, _lib.c_i32p, _lib.c_i32p, ctypes.c_int, _lib.c_i64p,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int, _lib.c_i32p, _lib.c_i64p,
                                           _lib.c_i32p, _lib.c_i32p, _lib.c_dp]
    L.pyhip_amg_plan_level_pfix.argtypes = [ctypes.c_void_p, ctypes.c_int, _lib.c_dp]
    L.pyhip_amg_plan_level_sizes.argtypes = [ctypes.c_void_p, ctypes.c_int, _lib.c_i64p]
    L.pyhip_amg_plan_level_copy.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_lib.c_i32p] * 12
    L.pyhip_amg_plan_free.argtypes = [ctypes.c_void_p]
    L.pyhip_amg_plan_error.restype = ctypes.c_char_p
    on, nc, frs, fcol, fw = fix
    arrs = [np.ascontiguousarray(x, t) for x, t in ((A.indptr, np.int32), (A.indices, np.int32), (off, np.int64),
                                                     (on, np.int32), (nc, np.int64), (frs, np.int32),
                                                     (fcol, np.int32), (fw, np.float64))]
    rs, col, so, on, nc, frs, fcol, fw = arrs
    h = L.pyhip_amg_plan_new_fixed(A.shape[0], rs.ctypes.data_as(_lib.c_i32p), col.ctypes.data_as(_lib.c_i32p),
                                   len(so) - 1, so.ctypes.data_as(_lib.c_i64p), cap, bottom, 12,
                                   on.ctypes.data_as(_lib.c_i32p), nc.ctypes.data_as(_lib.c_i64p),
                                   frs.ctypes.data_as(_lib.c_i32p), fcol.ctypes.data_as(_lib.c_i32p),
                                   fw.ctypes.data_as(_lib.c_dp))
    assert h, L.pyhip_amg_plan_error()
    sz = np.zeros(6, np.int64)
    L.pyhip_amg_plan_level_sizes(h, 0, sz.ctypes.data_as(_lib.c_i64p))
    n, ncoarse, nnz, nnzp, nnzq, ncp = (int(v) for v in sz)
    arr = dict(rs=n + 1, col=nnz, agg=n, prs=n + 1, pcol=nnzp, pcptr=nnzp + 1, pcidx=ncp, trs=ncoarse + 1,
               trow=nnzp, tsrc=nnzp, qrs=n + 1, qcol=nnzq)
    a = {k: np.zeros(max(v, 1), np.int32) for k, v in arr.items()}
    L.pyhip_amg_plan_level_copy(h, 0, *[a[k].ctypes.data_as(_lib.c_i32p) for k in arr])
    a = {k: a[k][:arr[k]] for k in arr}
    pfix = np.zeros(max(nnzp, 1))
    assert L.pyhip_amg_plan_level_pfix(h, 0, pfix.ctypes.data_as(_lib.c_dp)) == 0
    a.update(n=n, nc=ncoarse, pfix=pfix[:nnzp])
    L.pyhip_amg_plan_free(h)
    return a


def test_plan_with_the_element_level():
    """a P2 sub-problem between two P1 ones: its level-0 P is the interpolation (columns
    offset by the earlier sub-problems' coarse rows, weights as values before smoothing, no
    smoothing contributions), the P1 sub-problems aggregate as in the plain plan, and Q and
    the coarse pattern are the sparse products."""
    from pyoomph_amd import workloads as wl
    from tests.test_amg_plan import _plan
    parts = [wl.bratu(9, "C1", 1.0), wl.bratu(6, "C2", 1.0), wl.bratu(8, "C1", 1.0)]
    rows, cols, off = [], [], [0]
    fix_on, fix_nc, frs, fcol, fw = [], [], [0], [], []
    for mp in parts:
        eqn = mp.eqn[:, 0]
        E = eqn[mp.mesh.elements]
        k = E.shape[1]
        r, c = np.repeat(E, k, axis=1).ravel(), np.tile(E, (1, k)).ravel()
        m = (r >= 0) & (c >= 0)
        rows.append(r[m] + off[-1])
        cols.append(c[m] + off[-1])
        nc, rs, col, w = _rows(mp.mesh, eqn, mp.ndof)
        fix_on.append(1 if nc else 0)
        fix_nc.append(nc)
        frs += list(rs[1:] + frs[-1])
        fcol += list(col)
        fw += list(w)
        off.append(off[-1] + mp.ndof)
    n = off[-1]
    A = sp.csr_matrix((np.ones(sum(len(x) for x in rows)), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    lv = _plan_fixed(A, off, (fix_on, fix_nc, frs, fcol, fw))
    plain, _ = _plan(A.indptr, A.indices, off, cap=16, bottom=20)
    p0 = plain[0]
    # coarse offsets: sub 0 keeps its aggregates, sub 1 the vertices, sub 2 its aggregates
    agg0 = p0["agg"]
    na0 = len(np.unique(agg0[off[0]:off[1]]))
    na2 = len(np.unique(agg0[off[2]:off[3]]))
    assert lv["nc"] == na0 + fix_nc[1] + na2
    assert np.array_equal(lv["agg"][off[0]:off[1]], agg0[off[0]:off[1]])
    assert np.array_equal(lv["agg"][off[2]:off[3]] - (na0 + fix_nc[1]), agg0[off[2]:off[3]] - (agg0[off[2]:off[3]].min()))
    prs, pcol, pfix = lv["prs"], lv["pcol"], lv["pfix"]
    for r in range(n):
        q = slice(prs[r], prs[r + 1])
        if off[1] <= r < off[2]:
            t = slice(frs[r], frs[r + 1])
            assert np.array_equal(pcol[q], np.array(fcol[t], np.int32) + na0)
            assert np.array_equal(pfix[q], np.array(fw[t]))
            assert np.all(np.diff(lv["pcptr"])[q] == 0)
        else:
            assert np.array_equal(pfix[q], (pcol[q] == lv["agg"][r]).astype(float))
            assert np.all(np.diff(lv["pcptr"])[q] > 0)
    P = _bool(prs, pcol, (n, lv["nc"]))
    assert _same_pattern(P.T, lv["trs"], lv["trow"])
    assert _same_pattern(_bool(A.indptr, A.indices, A.shape) @ P, lv["qrs"], lv["qcol"])


def test_taylor_hood_velocity_components():
    """The block preconditioner's element level on Taylor-Hood (ux, uy at every Q2 node, p at
    the vertices): each velocity component is interpolated from its own vertex unknowns
    (coarse unknowns node-major: vertex k's ux, uy are 2k, 2k+1), reproducing bilinear
    functions; pressure rows are empty."""
    mesh = RectangularQuadMesh(4, space="C2")
    n = mesh.nnode
    verts = _vertices(mesh)
    eqn = -np.ones((n, 3), int)
    q = 0
    for node in range(n):
        for c in range(3):
            if c < 2 or node in set(verts):
                eqn[node, c] = q
                q += 1
    nc, rs, col, w = _rows(mesh, eqn.ravel(), q, form=_lib.FORM_NAVIER_STOKES_TH, nvalue=3)
    assert nc == 2 * len(verts)
    P = sp.csr_matrix((w, col, rs), shape=(q, nc))
    assert np.all(np.diff(rs)[eqn[verts, 2]] == 0)
    X = np.asarray(mesh.coords)
    rng = np.random.default_rng(4)
    for c in range(2):
        a = rng.uniform(-1, 1, 4)
        f = a[0] + a[1] * X[:, 0] + a[2] * X[:, 1] + a[3] * X[:, 0] * X[:, 1]
        coarse = np.zeros(nc)
        coarse[2 * np.arange(len(verts)) + c] = f[verts]
        assert np.max(np.abs((P @ coarse)[eqn[:, c]] - f)) < 1e-14
        assert np.max(np.abs((P @ coarse)[eqn[:, 1 - c]])) == 0.0
