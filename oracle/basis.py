"""Reference-element tables: multi-indices, Lagrange shape functions, quadrature.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""
import os
from itertools import combinations_with_replacement
from math import factorial

import numpy as np
from scipy.special import roots_jacobi


def _load_tables():
    # quadrature_rules.npz holds the rows of fealpy/quadrature/triangle.py:16-329
    # and fealpy/quadrature/tetrahedron.py:7-243, parsed from those files by
    # tests/golden/extract_reference_data.py (nothing here reads the product's copy).
    d = np.load(os.path.join(os.path.dirname(__file__), "quadrature_rules.npz"))
    tri = {int(k[4:]): d[k] for k in d.files if k.startswith("tri_")}
    tet = {int(k[4:]): d[k] for k in d.files if k.startswith("tet_")}
    return tri, tet


_TRI, _TET = _load_tables()


def multi_index_matrix(p, dim, dtype=np.int32):
    """fealpy/backend/numpy_backend.py:356-365."""
    sep = np.flip(np.array(tuple(combinations_with_replacement(range(p + 1), dim)),
                           dtype=dtype), axis=0)
    raw = np.zeros((sep.shape[0], dim + 2), dtype=dtype)
    raw[:, -1] = p
    raw[:, 1:-1] = sep
    return raw[:, 1:] - raw[:, :-1]


def simplex_shape_function(bc, p, mi=None):
    """fealpy/backend/numpy_backend.py:423-440."""
    if p == 1:
        return bc
    TD = bc.shape[-1] - 1
    if mi is None:
        mi = multi_index_matrix(p, TD)
    c = np.arange(1, p + 1, dtype=np.int_)
    P = 1.0 / np.multiply.accumulate(c)
    t = np.arange(0, p)
    shape = bc.shape[:-1] + (p + 1, TD + 1)
    A = np.ones(shape, dtype=bc.dtype)
    A[..., 1:, :] = p * bc[..., None, :] - t.reshape(-1, 1)
    np.cumprod(A, axis=-2, out=A)
    A[..., 1:, :] *= P.reshape(-1, 1)
    idx = np.arange(TD + 1)
    return np.prod(A[..., mi, idx], axis=-1)


def simplex_grad_shape_function(bc, p, mi=None):
    """fealpy/backend/numpy_backend.py:442-477: d phi_i / d lambda_m, (..., ldof, TD+1)."""
    TD = bc.shape[-1] - 1
    if mi is None:
        mi = multi_index_matrix(p, TD)
    ldof = mi.shape[0]
    c = np.arange(1, p + 1)
    P = 1.0 / np.multiply.accumulate(c)
    t = np.arange(0, p)
    shape = bc.shape[:-1] + (p + 1, TD + 1)
    A = np.ones(shape, dtype=bc.dtype)
    A[..., 1:, :] = p * bc[..., None, :] - t.reshape(-1, 1)
    FF = np.einsum('...jk, m->...kjm', A[..., 1:, :], np.ones(p))
    FF[..., range(p), range(p)] = p
    np.cumprod(FF, axis=-2, out=FF)
    F = np.zeros(shape, dtype=bc.dtype)
    F[..., 1:, :] = np.sum(np.tril(FF), axis=-1).swapaxes(-1, -2)
    F[..., 1:, :] *= P.reshape(-1, 1)
    np.cumprod(A, axis=-2, out=A)
    A[..., 1:, :] *= P.reshape(-1, 1)
    Q = A[..., mi, range(TD + 1)]
    M = F[..., mi, range(TD + 1)]
    R = np.zeros(bc.shape[:-1] + (ldof, TD + 1), dtype=bc.dtype)
    for i in range(TD + 1):
        idx = list(range(TD + 1))
        idx.remove(i)
        R[..., i] = M[..., i] * np.prod(Q[..., idx], axis=-1)
    return R


def stroud(dim, n):
    """fealpy/quadrature/stroud_quadrature.py:5-51 (conical Gauss-Jacobi product rule)."""
    points, weights = [], []
    for i in range(1, dim + 1):
        p, w, s = roots_jacobi(n, dim - i, 0, mu=True)
        points.append((p + 1) / 2)
        weights.append(w / s)
    points = np.meshgrid(*points)
    weights = np.meshgrid(*weights)
    pts = np.array([p.flatten() for p in points]).T
    ws = np.prod([w.flatten() for w in weights], axis=0)
    bcs = np.zeros(pts.shape[:-1] + (dim + 1,), dtype=np.float64)
    bcs[:, 0] = pts[:, 0]
    for i in range(1, dim):
        bcs[:, i] = pts[:, i] * (1 - bcs[:, :i].sum(axis=-1))
    bcs[:, dim] = 1 - bcs[:, :dim].sum(axis=-1)
    return bcs, ws


def quadrature(TD, q):
    """Cell rule used by ``mesh.quadrature_formula(q, 'cell')``.

    Triangle: fealpy/mesh/triangle_mesh.py:80-98 (q > 9 -> Stroud(2, q)).
    Tetrahedron: fealpy/mesh/tetrahedron_mesh.py:144-163 (q > 7 -> Stroud(3, q)).
    """
    if TD == 2:
        if q > 9:
            return stroud(2, q)
        A = np.array(_TRI[q], dtype=np.float64)
    elif TD == 3:
        if q > 7:
            return stroud(3, q)
        A = np.array(_TET[q], dtype=np.float64)
    else:
        raise ValueError(TD)
    return A[:, :-1], A[:, -1]


def factorial_int(n):
    return factorial(n)
