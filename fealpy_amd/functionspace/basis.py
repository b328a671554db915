"""Lagrange shape functions on the reference simplex (host side, FP64).

A degree-p Lagrange basis on a TD-simplex is indexed by multi-indices
alpha in N^{TD+1}, |alpha| = p, with
    phi_alpha(lambda) = prod_m  L_{alpha_m}(lambda_m),
    L_k(t) = prod_{s<k} (p t - s) / (s + 1),
so phi_alpha is 1 at the lattice point alpha/p and 0 at the others.  The
ordering of the multi-indices (first entry descending, then the rest
descending) is the DOF-local order of the reference
(fealpy/backend/numpy_backend.py:356-365), which fixes the meaning of every
column of cell_to_dof.  These tables are tiny; the kernels read them from HBM.
"""
from functools import lru_cache
from math import comb

import numpy as np


@lru_cache(maxsize=None)
def _multi_index(p, dim):
    out = []

    def rec(prefix, left, slots):
        if slots == 1:
            out.append(prefix + (left,))
            return
        for a in range(left, -1, -1):
            rec(prefix + (a,), left - a, slots - 1)

    rec((), p, dim + 1)
    mi = np.array(out, dtype=np.int32)
    assert mi.shape[0] == comb(p + dim, dim)
    mi.setflags(write=False)
    return mi


def multi_index_matrix(p, dim, dtype=np.int32):
    """(ldof, dim+1) multi-indices of the degree-p Lagrange points."""
    return _multi_index(p, dim).astype(dtype)


def _factor_tables(t, p):
    """L_k(t) and L_k'(t) for k = 0..p; t has shape (..., M)."""
    L = np.ones(t.shape + (p + 1,))
    dL = np.zeros(t.shape + (p + 1,))
    for k in range(1, p + 1):
        a = (p * t - (k - 1)) / k              # new linear factor
        dL[..., k] = dL[..., k - 1] * a + L[..., k - 1] * (p / k)
        L[..., k] = L[..., k - 1] * a
    return L, dL


def shape_function(bc, p, mi=None):
    """phi (..., ldof) at barycentric points bc (..., TD+1)."""
    bc = np.asarray(bc, dtype=np.float64)
    TD = bc.shape[-1] - 1
    mi = multi_index_matrix(p, TD) if mi is None else mi
    L, _ = _factor_tables(bc, p)              # (..., TD+1, p+1)
    cols = np.arange(TD + 1)
    return np.prod(L[..., cols, mi], axis=-1)  # (..., ldof)


def grad_shape_function(bc, p, mi=None):
    """d phi_alpha / d lambda_m, shape (..., ldof, TD+1)."""
    bc = np.asarray(bc, dtype=np.float64)
    TD = bc.shape[-1] - 1
    mi = multi_index_matrix(p, TD) if mi is None else mi
    L, dL = _factor_tables(bc, p)
    cols = np.arange(TD + 1)
    F = L[..., cols, mi]                       # (..., ldof, TD+1)
    dF = dL[..., cols, mi]
    out = np.empty(F.shape)
    for m in range(TD + 1):
        others = np.prod(np.delete(F, m, axis=-1), axis=-1)
        out[..., m] = dF[..., m] * others
    return out
