"""Host-side preparation of the linear-solver plugin (no GPU): the matrix built from
oomph's CSC-of-J^T arrays (unsorted, first-touch columns) equals J, and zero-diagonal
constraint rows get a structural diagonal plus the two-hop pattern."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import newton as on
from pyoomph_amd import _lib
from tests.oracle_bridge import oracle_problem


def build(n, rs, col, val, transpose=1):
    f = _lib.load().pyhip_plugin_build_csr_host
    f.restype = ctypes.c_int
    v = np.ascontiguousarray(val, dtype=np.float64)
    ri = np.ascontiguousarray(col, dtype=np.int32)
    cp = np.ascontiguousarray(rs, dtype=np.int32)
    cap = 8 * len(v) + n
    ors = np.zeros(n + 1, np.int32)
    ocl = np.zeros(cap, np.int32)
    osrc = np.zeros(cap, np.int64)
    m = f(n, len(v), v.ctypes.data_as(ctypes.c_void_p), ri.ctypes.data_as(ctypes.c_void_p),
          cp.ctypes.data_as(ctypes.c_void_p), transpose, ors.ctypes.data_as(ctypes.c_void_p),
          ocl.ctypes.data_as(ctypes.c_void_p), osrc.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(cap))
    assert m >= 0
    vals = np.where(osrc[:m] >= 0, v[np.maximum(osrc[:m], 0)], 0.0)
    return sp.csr_matrix((vals, ocl[:m], ors), shape=(n, n)), ors, ocl[:m], osrc[:m]


@pytest.mark.parametrize("kind,kw", [("poisson", dict(N=4)), ("ns", dict(N=3)),
                                     ("bratu", dict(N=5, space="C1", lam=1.0))])
def test_matrix_equals_jacobian(kind, kw):
    op = oracle_problem(kind, **kw)
    r, rs, col, val = op.get_jacobian(np.random.default_rng(0).uniform(-.3, .3, op.ndof))
    A, ors, ocl, src = build(op.ndof, rs, col, val)
    J = sp.csr_matrix((val, col, rs), shape=A.shape)
    assert abs(A - J).max() == 0.0
    for i in range(op.ndof):
        seg = ocl[ors[i]:ors[i + 1]]
        assert np.all(np.diff(seg) > 0) and i in seg        # sorted, diagonal present


def test_transpose_zero_reads_csc_of_a():
    rng = np.random.default_rng(1)
    D = rng.normal(size=(6, 6)) * (rng.random((6, 6)) < 0.5) + 4 * np.eye(6)
    C = sp.csc_matrix(D)
    A, *_ = build(6, C.indptr, C.indices, C.data, transpose=0)
    assert np.array_equal(A.toarray(), D)


def test_constraint_rows_get_two_hop_pattern():
    op = oracle_problem("ns", N=2)
    r, rs, col, val = op.get_jacobian(np.zeros(op.ndof))
    A, ors, ocl, src = build(op.ndof, rs, col, val)
    J = sp.csr_matrix((val, col, rs), shape=A.shape).toarray()
    zero = np.array([J[i, i] == 0 for i in range(op.ndof)])
    assert zero.any()
    S = (np.abs(J) > 0).astype(int)
    two_hop = (S[np.ix_(zero, ~zero)] @ S[np.ix_(~zero, zero)]) > 0
    added = (A.toarray() != 0) | (sp.csr_matrix((np.ones(len(ocl)), ocl, ors), shape=A.shape).toarray() > 0)
    assert np.all(added[np.ix_(zero, zero)][two_hop])
