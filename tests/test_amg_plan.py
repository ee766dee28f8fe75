"""Symbolic multigrid plan (host code, no GPU): every pattern checked against scipy's
sparse products on FE matrices of the oracle, batched over sub-problems."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

from pyoomph_amd import _lib
from tests.oracle_bridge import oracle_problem


def _plan(rs, col, sub_off, cap=12, bottom=64, maxlev=12):
    L = _lib.load()
    L.pyhip_amg_plan_new.restype = ctypes.c_void_p
    L.pyhip_amg_plan_new.argtypes = [ctypes.c_int64, _lib.c_i32p, _lib.c_i32p, ctypes.c_int, _lib.c_i64p,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.pyhip_amg_plan_levels.argtypes = [ctypes.c_void_p]
    L.pyhip_amg_plan_level_sizes.argtypes = [ctypes.c_void_p, ctypes.c_int, _lib.c_i64p]
    L.pyhip_amg_plan_level_copy.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_lib.c_i32p] * 12
    L.pyhip_amg_plan_bottom.argtypes = [ctypes.c_void_p, _lib.c_i64p, _lib.c_i32p, _lib.c_i32p, _lib.c_i64p]
    L.pyhip_amg_plan_free.argtypes = [ctypes.c_void_p]
    L.pyhip_amg_plan_error.restype = ctypes.c_char_p
    rs = np.ascontiguousarray(rs, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    so = np.ascontiguousarray(sub_off, np.int64)
    h = L.pyhip_amg_plan_new(len(rs) - 1, rs.ctypes.data_as(_lib.c_i32p), col.ctypes.data_as(_lib.c_i32p),
                             len(so) - 1, so.ctypes.data_as(_lib.c_i64p), cap, bottom, maxlev)
    assert h, L.pyhip_amg_plan_error()
    levels = []
    for l in range(L.pyhip_amg_plan_levels(h)):
        sz = np.zeros(6, np.int64)
        assert L.pyhip_amg_plan_level_sizes(h, l, sz.ctypes.data_as(_lib.c_i64p)) == 0
        n, nc, nnz, nnzp, nnzq, ncp = (int(v) for v in sz)
        arr = dict(rs=n + 1, col=nnz, agg=n, prs=n + 1, pcol=nnzp, pcptr=nnzp + 1, pcidx=ncp, trs=nc + 1,
                   trow=nnzp, tsrc=nnzp, qrs=n + 1, qcol=nnzq)
        a = {k: np.zeros(max(v, 1), np.int32) for k, v in arr.items()}
        assert L.pyhip_amg_plan_level_copy(h, l, *[a[k].ctypes.data_as(_lib.c_i32p) for k in arr]) == 0
        a = {k: a[k][:arr[k]] for k in arr}
        a.update(n=n, nc=nc)
        levels.append(a)
    sz = np.zeros(3, np.int64)
    L.pyhip_amg_plan_bottom(h, sz.ctypes.data_as(_lib.c_i64p), None, None, None)
    nb, nnzb, maxb = (int(v) for v in sz)
    brs = np.zeros(nb + 1, np.int32)
    bcol = np.zeros(max(nnzb, 1), np.int32)
    boff = np.zeros(len(so), np.int64)
    L.pyhip_amg_plan_bottom(h, sz.ctypes.data_as(_lib.c_i64p), brs.ctypes.data_as(_lib.c_i32p),
                            bcol.ctypes.data_as(_lib.c_i32p), boff.ctypes.data_as(_lib.c_i64p))
    L.pyhip_amg_plan_free(h)
    return levels, (brs, bcol[:nnzb], boff, maxb)


def _bool(rs, col, shape):
    return sp.csr_matrix((np.ones(len(col)), col, rs), shape=shape)


def _same_pattern(M, rs, col):
    M = M.tocsr()
    M.sum_duplicates()
    M.sort_indices()
    return np.array_equal(M.indptr, rs) and np.array_equal(M.indices, col)


def _batch(parts):
    """block-diagonal column-sorted structural pattern of several oracle problems"""
    mats, off = [], [0]
    for op in parts:
        G = op.geqn
        ns = G.shape[1]
        r = np.repeat(G, ns, axis=1).ravel()
        c = np.tile(G, (1, ns)).ravel()
        m = (r >= 0) & (c >= 0)
        S = sp.csr_matrix((np.ones(m.sum()), (r[m], c[m])), shape=(op.ndof, op.ndof))
        mats.append(S)
        off.append(off[-1] + op.ndof)
    A = sp.block_diag(mats).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return A, np.array(off)


@pytest.mark.parametrize("cap", [4, 12])
def test_plan_patterns_match_sparse_products(cap):
    parts = [oracle_problem("bratu", N=20, space="C2", lam=1.0), oracle_problem("bratu", N=9, space="C1", lam=2.0),
             oracle_problem("poisson", N=1), oracle_problem("rd", N=3, f=np.zeros(343))]
    A, off = _batch(parts)
    levels, (brs, bcol, boff, maxb) = _plan(A.indptr, A.indices, off, cap=cap, bottom=20)
    assert levels, "no coarsening"
    rs, col, sub_off = A.indptr, A.indices, off
    for lv in levels:
        n, nc = lv["n"], lv["nc"]
        assert np.array_equal(lv["rs"], rs) and np.array_equal(lv["col"], col)
        agg = lv["agg"]
        sizes_sub = np.diff(sub_off)
        cont = sizes_sub > 20
        ending = np.repeat(~cont, sizes_sub)
        # sub-problems at or below the bottom size end here: no aggregate, no P row
        assert np.all(agg[ending] == -1) and np.all(agg[~ending] >= 0)
        assert np.all(np.diff(lv["prs"])[ending] == 0)
        assert agg[~ending].max() == nc - 1 and len(np.unique(agg[~ending])) == nc
        sizes = np.bincount(agg[~ending], minlength=nc)
        assert sizes.max() <= cap + 1
        # aggregates stay inside sub-problems, coarse rows contiguous per sub-problem
        coff = [0]
        for s in range(len(sub_off) - 1):
            a = agg[sub_off[s]:sub_off[s + 1]]
            if cont[s]:
                assert a.min() == coff[-1] and a.max() - a.min() + 1 == len(np.unique(a))
                coff.append(coff[-1] + len(np.unique(a)))
            else:
                coff.append(coff[-1])
        Ab = _bool(rs, col, (n, n))
        Pt = sp.csr_matrix((np.ones(int((~ending).sum())), (np.nonzero(~ending)[0], agg[~ending])), shape=(n, nc))
        # P pattern = pattern of A P_t; contributions = the row's entries whose column is in J
        assert _same_pattern(Ab @ Pt, lv["prs"], lv["pcol"])
        for i in range(0, n, max(1, n // 97)):
            if ending[i]:
                continue
            for q in range(lv["prs"][i], lv["prs"][i + 1]):
                ks = lv["pcidx"][lv["pcptr"][q]:lv["pcptr"][q + 1]]
                want = [k for k in range(rs[i], rs[i + 1]) if agg[col[k]] == lv["pcol"][q]]
                assert list(ks) == want
        P = _bool(lv["prs"], lv["pcol"], (n, nc))
        # P^T rows list fine rows in increasing order, tsrc points back into P
        assert _same_pattern(P.T, lv["trs"], lv["trow"])
        rows_of_p = np.repeat(np.arange(n), np.diff(lv["prs"]))
        assert np.array_equal(rows_of_p[lv["tsrc"]], lv["trow"])
        assert np.array_equal(lv["pcol"][lv["tsrc"]], np.repeat(np.arange(nc), np.diff(lv["trs"])))
        Q = Ab @ P
        assert _same_pattern(Q, lv["qrs"], lv["qcol"])
        C = (P.T @ _bool(lv["qrs"], lv["qcol"], (n, nc))).tocsr()
        C.sum_duplicates()
        C.sort_indices()
        rs, col, sub_off = C.indptr, C.indices, np.array(coff)
    assert np.array_equal(brs, rs) and np.array_equal(bcol, col) and np.array_equal(boff, sub_off)
    assert maxb <= 20


def test_plan_of_a_small_problem_has_no_levels():
    A, off = _batch([oracle_problem("poisson", N=2)])
    levels, (brs, bcol, boff, maxb) = _plan(A.indptr, A.indices, off, bottom=64)
    assert levels == [] and np.array_equal(brs, A.indptr) and maxb == A.shape[0]


def test_each_sub_problem_gets_its_own_hierarchy():
    """the plan of a batch, restricted to one sub-problem, is that sub-problem's plan alone"""
    parts = [oracle_problem("bratu", N=20, space="C2", lam=1.0), oracle_problem("bratu", N=6, space="C2", lam=1.0),
             oracle_problem("bratu", N=13, space="C1", lam=1.0)]
    A, off = _batch(parts)
    batch, (_, _, boff, _) = _plan(A.indptr, A.indices, off, cap=12, bottom=30)
    for s in range(len(parts)):
        B = A[off[s]:off[s + 1], off[s]:off[s + 1]].tocsr()
        B.sort_indices()
        alone, (_, _, aoff, _) = _plan(B.indptr, B.indices, [0, B.shape[0]], cap=12, bottom=30)
        r0, r1 = off[s], off[s + 1]
        for l, lb in enumerate(batch):
            if l == len(alone):
                # the sub-problem's coarsest level: it ends here, with the rows it has alone
                assert r1 - r0 == aoff[1] - aoff[0]
                assert np.all(lb["agg"][r0:r1] == -1)
                r0 = r1 = 0
                continue
            if l > len(alone):
                break
            la = alone[l]
            a = lb["agg"][r0:r1]
            base = a[a >= 0].min() if (a >= 0).any() else 0
            assert np.array_equal(np.where(a >= 0, a - base, -1), la["agg"])
            assert np.array_equal(np.diff(lb["rs"][r0:r1 + 1]), np.diff(la["rs"]))
            # rows of this sub-problem on the next level
            nxt = np.unique(a[a >= 0])
            r0, r1 = (int(nxt.min()), int(nxt.max()) + 1) if len(nxt) else (0, 0)
        if len(alone) == len(batch):
            assert boff[s + 1] - boff[s] == aoff[1] - aoff[0]


def test_tiny_sub_problems_leave_empty_prolongation_blocks():
    """The batch of test_gpu_configs.test_multigrid_batch_with_empty_prolongation_blocks:
    the finest prolongation has 256-row blocks without entries, starting at an odd entry:
    every spmv variant must leave such a block without reading any entry."""
    from pyoomph_amd import workloads as wl
    specs = [(40, "C1", 1.0)] + [(7, "C1", 1.0 + 0.01 * i) for i in range(40)] + [(40, "C2", 1.2)]
    rows, cols, off = [], [], [0]
    for N, space, lam in specs:
        mp = wl.bratu(N, space, lam)
        E = mp.eqn[:, 0][mp.mesh.elements]
        k = E.shape[1]
        r, c = np.repeat(E, k, axis=1).ravel(), np.tile(E, (1, k)).ravel()
        m = (r >= 0) & (c >= 0)
        rows.append(r[m] + off[-1])
        cols.append(c[m] + off[-1])
        off.append(off[-1] + mp.ndof)
    n = off[-1]
    A = sp.csr_matrix((np.ones(sum(len(x) for x in rows)), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    A.sort_indices()
    levels, _ = _plan(A.indptr, A.indices, off, cap=16, bottom=48, maxlev=16)
    prs = levels[0]["prs"]
    empty_odd = [b for b in range((n + 255) // 256)
                 if prs[b * 256] == prs[min(n, (b + 1) * 256)] and prs[b * 256] % 2 == 1]
    assert len(empty_odd) >= 1
