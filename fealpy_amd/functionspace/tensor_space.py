"""Vector/tensor-valued Lagrange space (fealpy/functionspace/tensor_space.py:12-276).

``shape=(-1, k)`` puts the components of a scalar DOF next to each other
(global index k*scalar + component, ``dof_priority=False``); ``shape=(k, -1)``
stores component blocks one after the other (component*gdof + scalar,
``dof_priority=True``), as ``to_tensor_dof`` (functionspace/utils.py:65-95).
"""
from math import prod

import torch

from .lagrange import Function

_S = slice(None)


class TensorFunctionSpace:
    def __init__(self, scalar_space, shape):
        shape = tuple(shape)
        if len(shape) < 2:
            raise ValueError('shape must be a tuple of at least two element')
        if shape[0] == -1:
            self.dof_shape, self.dof_priority = shape[1:], False
        elif shape[-1] == -1:
            self.dof_shape, self.dof_priority = shape[:-1], True
        else:
            raise ValueError('`-1` is required as the first or last element')
        self.scalar_space = scalar_space
        self.shape = shape
        self.p = scalar_space.p
        self.TD = scalar_space.TD
        self.GD = scalar_space.GD
        self._c2d = None
        self._pattern = None
        self._bd = None

    @property
    def mesh(self):
        return self.scalar_space.mesh

    @property
    def device(self):
        return self.scalar_space.device

    @property
    def ftype(self):
        return self.scalar_space.ftype

    @property
    def itype(self):
        return self.scalar_space.itype

    @property
    def dof_numel(self):
        return prod(self.dof_shape)

    @property
    def dof_ndim(self):
        return len(self.dof_shape)

    def number_of_global_dofs(self):
        return self.dof_numel * self.scalar_space.number_of_global_dofs()

    def number_of_local_dofs(self, doftype='cell'):
        return self.dof_numel * self.scalar_space.number_of_local_dofs(doftype)

    def _expand(self, to_dof):
        k, g = self.dof_numel, self.scalar_space.number_of_global_dofs()
        to_dof = to_dof.to(torch.int64)
        comp = torch.arange(k, device=to_dof.device)
        if self.dof_priority:
            idx = comp[None, :, None] * g + to_dof[:, None, :]        # (entity, k, ldof)
        else:
            idx = to_dof[:, :, None] * k + comp[None, None, :]        # (entity, ldof, k)
        return idx.reshape(to_dof.shape[0], -1).to(torch.int32)

    def cell_to_dof(self, index=_S):
        if self._c2d is None:
            self._c2d = self._expand(self.scalar_space.cell_to_dof()).contiguous()
        return self._c2d[index]

    def face_to_dof(self, index=_S):
        return self._expand(self.scalar_space.face_to_dof())[index]

    def edge_to_dof(self, index=_S):
        return self._expand(self.scalar_space.edge_to_dof())[index]

    def interpolation_points(self):
        return self.scalar_space.interpolation_points()

    def _tensor_flags(self, scalar_flag):
        k = self.dof_numel
        if self.dof_priority:
            return scalar_flag[None, :].expand(k, -1).reshape(-1)
        return scalar_flag[:, None].expand(-1, k).reshape(-1)

    def _boundary(self):
        """(flags, indices) of boundary DOFs, cached; the scalar points come
        from the scalar space's cache."""
        if self._bd is None:
            flag = self._tensor_flags(self.scalar_space._boundary()[0]).contiguous()
            self._bd = (flag, torch.nonzero(flag)[:, 0])
        return self._bd

    def is_boundary_dof(self, threshold=None, method='interp'):
        """tensor_space.py:155-182."""
        if torch.is_tensor(threshold):
            if threshold.dtype == torch.bool and threshold.shape[0] == self.number_of_global_dofs():
                return threshold
            raise ValueError("len(threshold) must equal tensorspace gdof")
        if threshold is None:
            return self._boundary()[0].clone()
        if callable(threshold):
            return self._tensor_flags(self.scalar_space.is_boundary_dof(threshold, method=method))
        if isinstance(threshold, tuple):
            if len(threshold) != self.dof_numel:
                raise ValueError("one threshold per component is required")
            flags = [self.scalar_space.is_boundary_dof(t, method=method) for t in threshold]
            return torch.cat(flags) if self.dof_priority else torch.stack(flags, dim=1).reshape(-1)
        raise ValueError(f"Unknown type of threshold {type(threshold)}")

    def _layout(self, values):
        """(n_scalar, k) component values -> flat tensor-space order."""
        return values.T.reshape(-1) if self.dof_priority else values.reshape(-1)

    def boundary_interpolate(self, gd, uh=None, *, threshold=None, method=None):
        """uh[bd] = gd at the boundary interpolation points (tensor_space.py:184-276).
        Returns (uh, isTensorBDof)."""
        if uh is None:
            uh = torch.zeros(self.number_of_global_dofs(), dtype=self.ftype, device=self.device)
        arr = uh.array if isinstance(uh, Function) else uh
        k = self.dof_numel
        if isinstance(threshold, tuple):
            flag = self.is_boundary_dof(threshold, method=method)
            if callable(gd):
                vals = torch.as_tensor(gd(self.interpolation_points()), dtype=arr.dtype, device=arr.device)
                arr[flag] = self._layout(vals.reshape(-1, k))[flag]
            elif torch.is_tensor(gd):
                arr[flag] = gd[flag].to(arr.dtype)
            else:
                arr[flag] = gd
            return (uh if isinstance(uh, Function) else Function(self, arr)), flag
        if torch.is_tensor(threshold) or callable(threshold):
            flag = self.is_boundary_dof(threshold, method=method)
            if callable(gd):
                vals = torch.as_tensor(gd(self.interpolation_points()), dtype=arr.dtype, device=arr.device)
                arr[flag] = self._layout(vals.reshape(-1, k))[flag]
            elif torch.is_tensor(gd):
                arr[flag] = gd[flag].to(arr.dtype)
            else:
                arr[flag] = gd
            return (uh if isinstance(uh, Function) else Function(self, arr)), flag
        flag, idx = self._boundary()
        sflag, sidx, spts = self.scalar_space._boundary()
        if callable(gd):
            vals = torch.as_tensor(gd(spts), dtype=arr.dtype, device=arr.device).reshape(-1, k)   # (nbd, k)
            g = self.scalar_space.number_of_global_dofs()
            comp = torch.arange(k, device=arr.device)
            tidx = comp[None, :] * g + sidx[:, None] if self.dof_priority else sidx[:, None] * k + comp[None, :]
            arr[tidx.reshape(-1)] = vals.reshape(-1)
        elif torch.is_tensor(gd):
            arr[idx] = gd[idx].to(arr.dtype)
        elif isinstance(gd, (int, float)):
            arr[idx] = gd
        else:
            raise ValueError(f"Unknown type of gd {type(gd)}")
        return (uh if isinstance(uh, Function) else Function(self, arr)), flag.clone()

    set_dirichlet_bc = boundary_interpolate

    def function(self, array=None, dtype=None):
        return Function(self, array, dtype)

    def interpolate(self, u):
        """Values of u (returning (..., k)) at the interpolation points, in tensor-space order."""
        vals = torch.as_tensor(u(self.interpolation_points()), dtype=self.ftype, device=self.device)
        return Function(self, self._layout(vals.reshape(-1, self.dof_numel)).contiguous())

    def value(self, uh, bc, index=_S):
        """(NC, NQ, k) values at barycentric points."""
        k = self.dof_numel
        u = uh.reshape(k, -1).T if self.dof_priority else uh.reshape(-1, k)
        phi = self.scalar_space.basis(bc)[0].to(uh.dtype)               # (NQ, ldof)
        c2d = self.scalar_space.cell_to_dof(index).long()
        return torch.einsum('ql, cld -> cqd', phi, u[c2d])

    def tables(self, q):
        return self.scalar_space.tables(q)
