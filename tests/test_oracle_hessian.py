"""The oracle's Hessian-vector products, plain and transposed, pinned by central
differences of its Jacobian (which the element tests pin against finite differences of
the residual, and the known-answer tests against closed forms):

    plain       H v = d/de J(x + e v) y       (entry (i, k) = d (J y)_i / du_k)
    transposed  H v = d/de J(x + e v)^T y     (entry (i, k) = d (J^T y)_i / du_k)

The Navier-Stokes and reaction-diffusion Jacobians are polynomial of degree <= 2 in the
dofs, so the central difference is exact up to rounding there; Bratu's exponential leaves
an O(h^2) error. A transposed convection block, or a swapped row/column in the transposed
product, fails these checks."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle.assembly import map_order_csr
from tests.oracle_bridge import oracle_problem

CASES = [("ns", dict(N=3, Re=50.0), 1e-8), ("bratu", dict(N=4, space="C2", lam=1.5), 1e-6),
         ("bratu", dict(N=4, space="C1", lam=0.5), 1e-6), ("rd", dict(N=2), 1e-8)]


def _csr(op, rs, col, val):
    return sp.csr_matrix((val, col, rs), shape=(op.ndof, op.ndof))


def _jac(op, x):
    _, rs, col, val = op.get_jacobian(x)
    return _csr(op, rs, col, val)


def _kw(kind, kw):
    kw = dict(kw)
    if kind == "rd":
        from tests.oracle_bridge import product_problem
        kw["f"] = product_problem(kind, **kw).equations.source_field
    return kw


@pytest.mark.parametrize("kind,kw,tol", CASES)
@pytest.mark.parametrize("transposed", [False, True])
def test_hessian_products_are_derivatives_of_the_jacobian(kind, kw, tol, transposed):
    op = oracle_problem(kind, **_kw(kind, kw))
    n = op.ndof
    rng = np.random.default_rng(5)
    x, y, v = rng.uniform(-0.4, 0.4, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    _, rs, col, val = op.get_hessian_vector_product(x, y, transposed=transposed)
    H = _csr(op, rs, col, val)
    h = 1e-4
    Jp, Jm = _jac(op, x + h * v), _jac(op, x - h * v)
    fd = ((Jp.T @ y - Jm.T @ y) if transposed else (Jp @ y - Jm @ y)) / (2 * h)
    assert np.max(np.abs(H @ v - fd)) <= tol * np.max(np.abs(fd))


def test_transposed_product_of_navier_stokes_is_the_symmetric_hessian_of_y_dot_r():
    """Transposed: the Hessian of the scalar y . R(u), so symmetric; the plain product is
    not (the convection term is not a gradient)."""
    op = oracle_problem("ns", N=3, Re=50.0)
    rng = np.random.default_rng(9)
    x, y = rng.uniform(-0.4, 0.4, op.ndof), rng.uniform(-1, 1, op.ndof)
    Ht = _csr(op, *op.get_hessian_vector_product(x, y, transposed=True)[1:])
    Hp = _csr(op, *op.get_hessian_vector_product(x, y)[1:])
    scale = np.max(np.abs(Ht.toarray()))
    assert np.max(np.abs((Ht - Ht.T).toarray())) <= 1e-13 * scale
    assert np.max(np.abs((Hp - Hp.T).toarray())) > 1e-3 * scale


@pytest.mark.parametrize("kind,kw", [("bratu", dict(N=4, space="C2", lam=1.5)), ("rd", dict(N=2))])
def test_transposed_product_of_scalar_forms_is_the_plain_one(kind, kw):
    """d^2 R_i/(du_j du_k) = int d2src psi_i psi_j psi_k is symmetric in i, j, k."""
    op = oracle_problem(kind, **_kw(kind, kw))
    rng = np.random.default_rng(2)
    x, y = rng.uniform(-0.4, 0.4, op.ndof), rng.uniform(-1, 1, op.ndof)
    a = op.get_hessian_vector_product(x, y, transposed=True)
    b = op.get_hessian_vector_product(x, y)
    for u, w in zip(a, b):
        assert np.array_equal(u, w)


def test_map_order_csr_sorts_rows_and_keeps_entries():
    rs = np.array([0, 3, 3, 5])
    col = np.array([4, 0, 2, 1, 0])
    val = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    r2, c2, v2 = map_order_csr(rs, col, val)
    assert np.array_equal(r2, rs)
    assert np.array_equal(c2, [0, 2, 4, 0, 1]) and np.array_equal(v2, [2.0, 3.0, 1.0, 5.0, 4.0])
