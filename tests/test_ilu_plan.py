"""ILU(0) row order and level schedule (host code, no GPU): the permuted pattern is the
symmetric permutation of the matrix pattern, every level of the triangular solves holds
rows that do not depend on each other, the multicolour order keeps same-coloured rows of
a group apart, and each sub-problem's order is the one it has alone in the batch."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

from pyoomph_amd import _lib
from tests.oracle_bridge import oracle_problem

AUTO, NATURAL, MULTICOLOR = 0, 1, 2


def _lib_ilu():
    L = _lib.load()
    L.pyhip_ilu_plan_new.restype = ctypes.c_void_p
    L.pyhip_ilu_plan_new.argtypes = [ctypes.c_int64, _lib.c_i32p, _lib.c_i32p, ctypes.c_int, _lib.c_i64p,
                                     ctypes.POINTER(ctypes.c_int8), ctypes.c_int, ctypes.c_int]
    L.pyhip_ilu_plan_sizes.argtypes = [ctypes.c_void_p, _lib.c_i64p]
    L.pyhip_ilu_plan_copy.argtypes = [ctypes.c_void_p] + [_lib.c_i32p] * 6 + [_lib.c_i64p, _lib.c_i32p, _lib.c_i64p]
    L.pyhip_ilu_plan_free.argtypes = [ctypes.c_void_p]
    L.pyhip_ilu_plan_error.restype = ctypes.c_char_p
    return L


def _plan(rs, col, sub_off, is_p=None, order=AUTO, auto_levels=64):
    L = _lib_ilu()
    rs = np.ascontiguousarray(rs, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    so = np.ascontiguousarray(sub_off, np.int64)
    n = len(rs) - 1
    g = None if is_p is None else np.ascontiguousarray(is_p, np.int8)
    h = L.pyhip_ilu_plan_new(n, rs.ctypes.data_as(_lib.c_i32p), col.ctypes.data_as(_lib.c_i32p), len(so) - 1,
                             so.ctypes.data_as(_lib.c_i64p),
                             None if g is None else g.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)), order,
                             auto_levels)
    if not h:
        raise RuntimeError(L.pyhip_ilu_plan_error().decode())
    sz = np.zeros(4, np.int64)
    assert L.pyhip_ilu_plan_sizes(h, sz.ctypes.data_as(_lib.c_i64p)) == 0
    nlo, nup, nnz, mc = (int(v) for v in sz)
    out = dict(perm=np.zeros(n, np.int32), prs=np.zeros(n + 1, np.int32), pcol=np.zeros(nnz, np.int32),
               psrc=np.zeros(nnz, np.int32), pdiag=np.zeros(n, np.int32), lo_rows=np.zeros(n, np.int32),
               lo_ptr=np.zeros(nlo + 1, np.int64), up_rows=np.zeros(n, np.int32), up_ptr=np.zeros(nup + 1, np.int64))
    ptr = [v.ctypes.data_as(_lib.c_i64p if v.dtype == np.int64 else _lib.c_i32p) for v in out.values()]
    assert L.pyhip_ilu_plan_copy(h, *ptr) == 0
    L.pyhip_ilu_plan_free(h)
    out["multicolor"] = bool(mc)
    return out


def _batch(parts):
    """block-diagonal column-sorted structural pattern of several oracle problems, with
    the group of each row (1: Taylor-Hood pressure)"""
    mats, off, grp = [], [0], []
    for op in parts:
        G = op.geqn
        ns = G.shape[1]
        r = np.repeat(G, ns, axis=1).ravel()
        c = np.tile(G, (1, ns)).ravel()
        m = (r >= 0) & (c >= 0)
        mats.append(sp.csr_matrix((np.ones(m.sum()), (r[m], c[m])), shape=(op.ndof, op.ndof)))
        off.append(off[-1] + op.ndof)
        g = np.zeros(op.ndof, np.int8)
        if op.nvalue == 3:
            p = op.eqn[:, 2]
            g[p[p >= 0]] = 1
        grp.append(g)
    A = sp.block_diag(mats).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return A, np.array(off), np.concatenate(grp)


def _check_plan(A, off, grp, P):
    n = A.shape[0]
    rs, col = A.indptr, A.indices
    perm = P["perm"]
    assert np.array_equal(np.sort(perm), np.arange(n))
    iperm = np.empty(n, np.int64)
    iperm[perm] = np.arange(n)
    # each sub-problem keeps its own block of positions; group 0 rows come first
    for s in range(len(off) - 1):
        blk = perm[off[s]:off[s + 1]]
        assert blk.min() == off[s] and blk.max() == off[s + 1] - 1
        g = grp[blk]
        assert np.all(np.diff(g.astype(int)) >= 0)
    # permuted pattern = pattern of Pi A Pi^T, column-sorted, psrc points to the entry
    B = A[perm][:, perm].tocsr()
    B.sort_indices()
    assert np.array_equal(P["prs"], B.indptr) and np.array_equal(P["pcol"], B.indices)
    rows_new = np.repeat(np.arange(n), np.diff(P["prs"]))
    rows_old = np.repeat(np.arange(n), np.diff(rs))
    assert np.array_equal(rows_old[P["psrc"]], perm[rows_new])
    assert np.array_equal(col[P["psrc"]], perm[P["pcol"]])
    assert np.array_equal(P["pcol"][P["pdiag"]], np.arange(n))
    # level sets: every row in exactly one level; a row's lower (upper) neighbours sit in
    # earlier levels, so the rows of one level do not depend on each other
    for rows, ptr, side in ((P["lo_rows"], P["lo_ptr"], -1), (P["up_rows"], P["up_ptr"], 1)):
        assert ptr[0] == 0 and ptr[-1] == n and np.all(np.diff(ptr) > 0)
        assert np.array_equal(np.sort(rows), np.arange(n))
        lev = np.empty(n, np.int64)
        lev[rows] = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
        pc = P["pcol"]
        dep = (pc < rows_new) if side < 0 else (pc > rows_new)
        assert np.all(lev[pc[dep]] < lev[rows_new[dep]])
        # each level is as early as it can be (one more than its deepest dependency)
        deepest = np.full(n, -1, np.int64)
        np.maximum.at(deepest, rows_new[dep], lev[pc[dep]])
        assert np.array_equal(lev, deepest + 1)
    return iperm


def test_natural_order_is_the_identity_and_levels_follow_the_mesh():
    parts = [oracle_problem("bratu", N=9, space="C2", lam=1.0), oracle_problem("poisson", N=2)]
    A, off, grp = _batch(parts)
    P = _plan(A.indptr, A.indices, off, order=NATURAL)
    assert np.array_equal(P["perm"], np.arange(A.shape[0])) and not P["multicolor"]
    _check_plan(A, off, grp, P)


@pytest.mark.parametrize("order", [NATURAL, MULTICOLOR])
def test_taylor_hood_pressures_follow_velocities(order):
    parts = [oracle_problem("ns", N=6), oracle_problem("bratu", N=5, space="C1", lam=1.0), oracle_problem("ns", N=4)]
    A, off, grp = _batch(parts)
    assert grp.sum() > 0
    P = _plan(A.indptr, A.indices, off, is_p=grp, order=order)
    _check_plan(A, off, grp, P)


def test_multicolour_order_separates_neighbours():
    parts = [oracle_problem("bratu", N=16, space="C2", lam=1.0), oracle_problem("ns", N=6),
             oracle_problem("rd", N=4, f=np.zeros(729))]
    A, off, grp = _batch(parts)
    P = _plan(A.indptr, A.indices, off, is_p=grp, order=MULTICOLOR)
    assert P["multicolor"]
    _check_plan(A, off, grp, P)
    # rows of one colour are a run of consecutive positions with no coupling between them;
    # the colour classes are the lower-solve levels restricted to one group
    n = A.shape[0]
    lev = np.empty(n, np.int64)
    lev[P["lo_rows"]] = np.repeat(np.arange(len(P["lo_ptr"]) - 1), np.diff(P["lo_ptr"]))
    rows_new = np.repeat(np.arange(n), np.diff(P["prs"]))
    pc = P["pcol"]
    g_new = grp[P["perm"]]
    same = (g_new[rows_new] == g_new[pc]) & (pc != rows_new)
    assert np.all(lev[rows_new[same]] != lev[pc[same]])
    # greedy colouring needs at most (largest group degree + 1) colours per group, so the
    # lower solve of the batch has at most that many levels per group
    deg = np.bincount(rows_new[same], minlength=n).max()
    assert len(P["lo_ptr"]) - 1 <= 2 * (deg + 1)
    Pn = _plan(A.indptr, A.indices, off, is_p=grp, order=NATURAL)
    assert len(P["lo_ptr"]) < len(Pn["lo_ptr"])


def _restrict(P, off, s):
    """the plan of sub-problem s, in its own numbering"""
    r0, r1 = off[s], off[s + 1]
    perm = P["perm"][r0:r1] - r0
    prs = P["prs"][r0:r1 + 1] - P["prs"][r0]
    pcol = P["pcol"][P["prs"][r0]:P["prs"][r1]] - r0
    return perm, prs, pcol


def test_each_sub_problem_keeps_its_own_order_in_a_batch():
    """auto order: a sub-problem switches to multicolour by its own level count, so its
    order is the same alone and in any batch"""
    parts = [oracle_problem("bratu", N=24, space="C2", lam=1.0), oracle_problem("bratu", N=3, space="C2", lam=1.0),
             oracle_problem("ns", N=8), oracle_problem("poisson", N=1)]
    A, off, grp = _batch(parts)
    P = _plan(A.indptr, A.indices, off, is_p=grp, order=AUTO, auto_levels=20)
    _check_plan(A, off, grp, P)
    switched = []
    for s in range(len(parts)):
        B = A[off[s]:off[s + 1], off[s]:off[s + 1]].tocsr()
        B.sort_indices()
        g = grp[off[s]:off[s + 1]]
        alone = _plan(B.indptr, B.indices, [0, B.shape[0]], is_p=g, order=AUTO, auto_levels=20)
        perm, prs, pcol = _restrict(P, off, s)
        assert np.array_equal(perm, alone["perm"])
        assert np.array_equal(prs, alone["prs"]) and np.array_equal(pcol, alone["pcol"])
        switched.append(alone["multicolor"])
    # the large sub-problems switch, the small ones keep the natural order
    assert switched[0] and not switched[1] and not switched[3]
    assert P["multicolor"]


def test_missing_diagonal_is_an_error():
    A = sp.csr_matrix(np.array([[1.0, 1.0, 0.0], [1.0, 0.0, 1.0], [0.0, 1.0, 1.0]]))
    A.eliminate_zeros()
    with pytest.raises(RuntimeError, match="diagonal"):
        _plan(A.indptr, A.indices, [0, 3], order=NATURAL)


def test_level_schedule_solves_like_a_sequential_sweep():
    """a triangular solve done level by level (all rows of a level at once, from values
    of earlier levels) equals the row-by-row forward and backward substitution"""
    parts = [oracle_problem("ns", N=5), oracle_problem("bratu", N=7, space="C2", lam=1.0)]
    A, off, grp = _batch(parts)
    P = _plan(A.indptr, A.indices, off, is_p=grp, order=MULTICOLOR)
    n = A.shape[0]
    rng = np.random.default_rng(3)
    B = A[P["perm"]][:, P["perm"]].tocsr()
    B.sort_indices()
    B.data = rng.uniform(0.1, 1.0, B.nnz)
    B.setdiag(B.diagonal() + 10.0)
    b = rng.standard_normal(n)
    Lo = sp.tril(B, -1).tocsr()
    y = np.zeros(n)
    for l in range(len(P["lo_ptr"]) - 1):
        rows = P["lo_rows"][P["lo_ptr"][l]:P["lo_ptr"][l + 1]]
        y[rows] = b[rows] - Lo[rows] @ y
    ys = np.zeros(n)
    for i in range(n):
        ys[i] = b[i] - (Lo[i] @ ys)[0]
    assert np.allclose(y, ys, rtol=1e-13, atol=1e-13)
    x = np.zeros(n)
    d = B.diagonal()
    Us = sp.triu(B, 1).tocsr()
    for l in range(len(P["up_ptr"]) - 1):
        rows = P["up_rows"][P["up_ptr"][l]:P["up_ptr"][l + 1]]
        x[rows] = (y[rows] - Us[rows] @ x) / d[rows]
    xs = np.zeros(n)
    for i in range(n - 1, -1, -1):
        xs[i] = (y[i] - (Us[i] @ xs)[0]) / d[i]
    assert np.allclose(x, xs, rtol=1e-13, atol=1e-13)
