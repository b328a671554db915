"""Vector integrators: LinearElasticityIntegrator
(fealpy/fem/linear_elasticity_integrator.py:14-186) and VectorSourceIntegrator
(fealpy/fem/vector_source_integrator.py:12-52), on a TensorFunctionSpace over
a scalar Lagrange space.  Element tensors come from fa_element_elasticity /
fa_element_vectors_tensor; BilinearForm/LinearForm gather them like any
other integrator."""
import torch

from .. import _native as nat
from .integrators import Integrator

_S = slice(None)


def _check_tensor_space(space, who):
    if not hasattr(space, 'scalar_space') or space.dof_ndim != 1 or space.dof_numel != space.GD:
        raise RuntimeError(f"{who} needs a TensorFunctionSpace with shape (-1, GD) or (GD, -1) over a "
                           f"Lagrange space")


def _precontracted(space, q):
    """M[i, j, m, n] = sum_q w_q R[q, i, m] R[q, j, n] (FP64, on the device),
    cached per quadrature index."""
    ss = space.scalar_space
    cache = ss.__dict__.setdefault('_elast_M', {})
    if q not in cache:
        _, qw, _, R = ss.tables(q)                                   # R: (NQ, ldof, TD+1)
        cache[q] = torch.einsum('q, qim, qjn -> ijmn', qw, R, R).contiguous()
    return cache[q]


class LinearElasticityIntegrator(Integrator):
    """Isotropic linear elasticity on simplices.  The element tensor has the
    block structure of linear_elasticity_integrator.py:90-165 for either DOF
    layout of the tensor space."""

    def __init__(self, material, q=None, *, index=_S, method=None):
        super().__init__()
        self.material, self.q, self.index, self.method = material, q, index, method
        if not (isinstance(index, slice) and index == _S):
            self.set_region(index)

    def to_global_dof(self, space, /, indices=None):
        return space.cell_to_dof(index=self.entity_selection(indices))

    def _coefficients(self, GD, dof_priority):
        D = self.material.elastic_matrix()
        if D.shape[1] != 1 or D.shape[0] != 1:
            raise ValueError("assembly currently only supports elastic matrices of shape (1, 1, n, n)")
        D = D[0, 0]
        if GD == 3:
            d00, d01, d55 = float(D[0, 0]), float(D[0, 1]), float(D[5, 5])
            cA = d00 if dof_priority else 2 * d55 + d01
            return cA, d55, d01, d55
        d00, d01, d22 = float(D[0, 0]), float(D[0, 1]), float(D[2, 2])
        return d00, d22, d01, d22

    def assembly(self, space, /, indices=None, dtype=None):
        _check_tensor_space(space, "LinearElasticityIntegrator")
        ss = space.scalar_space
        mesh = ss.mesh
        sel = self.entity_selection(indices)
        q = ss.p + 3 if self.q is None else self.q
        M = _precontracted(space, q)
        cell = mesh.cell[sel].to(torch.int32).contiguous()
        nc = cell.shape[0]
        ldof = ss.number_of_local_dofs()
        GD = space.GD
        cA, cB, cC, cD = self._coefficients(GD, space.dof_priority)
        dtype = torch.float64 if dtype is None else dtype
        elem = torch.empty((nc, GD * ldof, GD * ldof), dtype=dtype, device=space.device)
        nat.check(nat.lib().fa_element_elasticity(nat.dtype_code(dtype), ss.TD, ldof, nat.ptr(mesh.node),
                                                  nat.ptr(cell), nc, nat.ptr(M), cA, cB, cC, cD,
                                                  int(space.dof_priority), nat.ptr(elem), nat.stream()),
                  "fa_element_elasticity")
        return elem


class VectorSourceIntegrator(Integrator):
    """f . v for a vector source: a callable returning (NC, NQ, GD) (cartesian
    or barycentric), a (NC, NQ, GD) / (NC, GD) / (GD,) tensor or a number."""

    def __init__(self, source=None, q=None, *, index=_S, batched=False):
        super().__init__()
        self.source, self.q, self.index, self.batched = source, q, index, batched
        if not (isinstance(index, slice) and index == _S):
            self.set_region(index)

    def to_global_dof(self, space, /, indices=None):
        return space.cell_to_dof(index=self.entity_selection(indices))

    def assembly(self, space, /, indices=None, dtype=None):
        if self.batched:
            raise NotImplementedError("batched sources are not provided")
        _check_tensor_space(space, "VectorSourceIntegrator")
        ss = space.scalar_space
        mesh = ss.mesh
        sel = self.entity_selection(indices)
        q = ss.p + 3 if self.q is None else self.q
        bcs, qw, phi, _ = ss.tables(q)
        cell = mesh.cell[sel].to(torch.int32).contiguous()
        nc, nq, k = cell.shape[0], qw.shape[0], space.dof_numel
        f = self.source
        if f is None:
            f = 1.0
        if callable(f) and not torch.is_tensor(f):
            if getattr(f, 'coordtype', 'barycentric') == 'barycentric':
                f = f(bcs.to(space.device), index=sel)
            else:
                f = f(mesh.bc_to_point(bcs, index=sel))
        f = torch.as_tensor(f, dtype=torch.float64, device=space.device)
        if f.ndim == 0:
            f = f.expand(nc, nq, k)
        elif f.ndim == 1 and f.shape[0] == k:
            f = f[None, None, :].expand(nc, nq, k)
        elif f.ndim == 2 and tuple(f.shape) == (nc, k):
            f = f[:, None, :].expand(nc, nq, k)
        if tuple(f.shape) != (nc, nq, k):
            raise ValueError(f"vector source of shape {tuple(f.shape)} is not supported "
                             f"(expected (NC, NQ, {k}), (NC, {k}), ({k},) or a number)")
        f = f.contiguous()
        ldof = ss.number_of_local_dofs()
        dtype = torch.float64 if dtype is None else dtype
        elem = torch.empty((nc, ldof * k), dtype=dtype, device=space.device)
        nat.check(nat.lib().fa_element_vectors_tensor(nat.dtype_code(dtype), ss.TD, ldof, k,
                                                      int(space.dof_priority), nq, nat.ptr(mesh.node),
                                                      nat.ptr(cell), nc, nat.ptr(qw), nat.ptr(phi), nat.ptr(f),
                                                      nat.ptr(elem), nat.stream()), "fa_element_vectors_tensor")
        return elem
