"""Quadrature rules on simplices, as selected by ``mesh.quadrature_formula``.

``q`` is a rule *index*, not a degree (fealpy/mesh/triangle_mesh.py:80-98,
fealpy/mesh/tetrahedron_mesh.py:144-163): tabulated symmetric rules up to
index 9 (triangle) / 7 (tetrahedron), the conical Stroud rule above that.
Points and weights are FP64 torch tensors (barycentric points (NQ, TD+1)).
"""
import numpy as np
import torch
from scipy.special import roots_jacobi

from .tables import TETRAHEDRON_RULES, TRIANGLE_RULES

__all__ = ["Quadrature", "TriangleQuadrature", "TetrahedronQuadrature", "StroudQuadrature", "simplex_rule"]


class Quadrature:
    """Base: holds ``quadpts`` (NQ, TD+1) and ``weights`` (NQ,)."""

    def __init__(self, points, weights, device=None):
        self.quadpts = torch.as_tensor(points, dtype=torch.float64, device=device)
        self.weights = torch.as_tensor(weights, dtype=torch.float64, device=device)

    def number_of_quadrature_points(self):
        return self.weights.shape[0]

    __len__ = number_of_quadrature_points

    def get_quadrature_points_and_weights(self):
        return self.quadpts, self.weights

    def get_quadrature_point_and_weight(self, i):
        return self.quadpts[i, :], self.weights[i]


class _Tabulated(Quadrature):
    TABLE = {}

    def __init__(self, index, *, dtype=None, device=None):
        if index not in self.TABLE:
            raise NotImplementedError(f"quadrature index {index} is not tabulated "
                                      f"(available {min(self.TABLE)}..{max(self.TABLE)})")
        A = np.array(self.TABLE[index], dtype=np.float64)
        super().__init__(A[:, :-1], A[:, -1], device)


class TriangleQuadrature(_Tabulated):
    """fealpy/quadrature/triangle.py:15-329."""
    TABLE = TRIANGLE_RULES


class TetrahedronQuadrature(_Tabulated):
    """fealpy/quadrature/tetrahedron.py:6-243."""
    TABLE = TETRAHEDRON_RULES


class StroudQuadrature(Quadrature):
    """Collapsed (conical) product rule of Stroud on the d-simplex, exact to
    degree 2n-1: Gauss-Jacobi nodes in each collapsed coordinate
    (fealpy/quadrature/stroud_quadrature.py:5-51)."""

    def __init__(self, dim, n, device=None):
        self.dim, self.n = dim, n
        axes_p, axes_w = [], []
        for i in range(1, dim + 1):
            x, w, mu = roots_jacobi(n, dim - i, 0, mu=True)
            axes_p.append((x + 1) / 2)
            axes_w.append(w / mu)
        P = np.meshgrid(*axes_p)
        W = np.meshgrid(*axes_w)
        t = np.stack([a.ravel() for a in P], axis=1)          # collapsed coordinates
        w = np.prod(np.stack([a.ravel() for a in W], axis=0), axis=0)
        bc = np.empty((t.shape[0], dim + 1))
        rest = np.ones(t.shape[0])
        bc[:, 0] = t[:, 0]
        for i in range(1, dim):
            bc[:, i] = t[:, i] * (1 - bc[:, :i].sum(axis=1))
        bc[:, dim] = 1 - bc[:, :dim].sum(axis=1)
        del rest
        super().__init__(bc, w, device)


def simplex_rule(TD, q, device=None):
    """The cell rule ``mesh.quadrature_formula(q, 'cell')`` returns."""
    if TD == 2:
        return StroudQuadrature(2, q, device) if q > 9 else TriangleQuadrature(q, device=device)
    if TD == 3:
        return StroudQuadrature(3, q, device) if q > 7 else TetrahedronQuadrature(q, device=device)
    raise ValueError(f"unsupported topological dimension {TD}")
