"""Continuous Lagrange space on a simplex mesh (fealpy/functionspace/lagrange_fe_space.py)."""
from math import comb

import numpy as np
import torch

from .basis import grad_shape_function, multi_index_matrix, shape_function

_S = slice(None)


def _as_points_array(bcs):
    return bcs.detach().cpu().numpy() if torch.is_tensor(bcs) else np.asarray(bcs, dtype=np.float64)


class Function:
    """A finite-element function: coefficient vector + space
    (fealpy/functionspace/function.py:12-94).  Calling it evaluates it at
    barycentric points on every cell: (NC, NQ)."""
    coordtype = 'barycentric'

    def __init__(self, space, array=None, dtype=None):
        self.space = space
        if array is None:
            array = torch.zeros(space.number_of_global_dofs(), dtype=dtype or space.ftype, device=space.device)
        self.array = array

    def __getitem__(self, idx):
        return self.array[idx]

    def __setitem__(self, idx, val):
        self.array[idx] = torch.as_tensor(val, dtype=self.array.dtype, device=self.array.device)

    def __call__(self, bcs, index=_S):
        return self.space.value(self.array, bcs, index)

    @property
    def grad_value(self):
        """Barycentric callable bcs -> (NC, NQ, GD) gradient of the function
        (function.py's attribute forwarding of space.grad_value)."""
        def gv(bcs, index=_S):
            return self.space.grad_value(self.array, bcs, index)
        gv.coordtype = 'barycentric'
        return gv

    def value(self, bcs, index=_S):
        return self.space.value(self.array, bcs, index)

    def __len__(self):
        return self.array.shape[0]

    @property
    def shape(self):
        return self.array.shape

    def numpy(self):
        return self.array.detach().cpu().numpy()


class LagrangeFESpace:
    """Degree-p continuous Lagrange space; DOF numbering from the mesh
    (``cell_to_ipoint``), FP64 geometry, tables cached per (p, q)."""

    def __init__(self, mesh, p=1, ctype='C'):
        if ctype != 'C':
            raise NotImplementedError("only continuous ('C') Lagrange spaces are provided")
        if not 1 <= p <= 4:
            raise ValueError("p must be in 1..4")
        self.mesh = mesh
        self.p = p
        self.ctype = ctype
        self.ftype = mesh.ftype
        self.itype = mesh.itype
        self.device = mesh.device
        self.TD = mesh.top_dimension()
        self.GD = mesh.geo_dimension()
        self._c2d = None
        self._tables = {}
        self._pattern = None
        self._bd = None

    def __str__(self):
        return "Lagrange finite element space on linear mesh!"

    def number_of_local_dofs(self, doftype='cell'):
        return self.mesh.number_of_local_ipoints(self.p, doftype)

    def number_of_global_dofs(self):
        return self.mesh.number_of_global_ipoints(self.p)

    def cell_to_dof(self, index=_S):
        if self._c2d is None:
            self._c2d = self.mesh.cell_to_ipoint(self.p).to(torch.int32).contiguous()
        return self._c2d[index]

    def face_to_dof(self, index=_S):
        return self.mesh.face_to_ipoint(self.p)[index]

    def edge_to_dof(self, index=_S):
        return self.mesh.edge_to_ipoint(self.p)[index]

    def interpolation_points(self):
        return self.mesh.interpolation_points(self.p)

    def geo_dimension(self):
        return self.GD

    def top_dimension(self):
        return self.TD

    def _boundary(self):
        """(flags, indices, interpolation points) of the DOFs on boundary faces,
        computed once per space: Dirichlet data is re-imposed on every assembly
        and these depend only on the mesh."""
        if self._bd is None:
            flag = self.mesh.is_boundary_dof(self.p)
            idx = torch.nonzero(flag)[:, 0]
            self._bd = (flag, idx, self.interpolation_points()[idx].contiguous())
        return self._bd

    def is_boundary_dof(self, threshold=None, method=None):
        """DOFs on boundary faces (fealpy/functionspace/dofs.py:23-55)."""
        gdof = self.number_of_global_dofs()
        if threshold is None and self._bd is not None:
            return self._bd[0].clone()
        if torch.is_tensor(threshold):
            if threshold.dtype == torch.bool and threshold.shape[0] == gdof:
                return threshold
            raise ValueError(f"Unknown threshold: {threshold}")
        mesh = self.mesh
        index = mesh.boundary_face_index()
        flag = torch.zeros(gdof, dtype=torch.bool, device=self.device)
        if method in (None, 'centroid'):
            if callable(threshold):
                index = index[threshold(mesh.entity_barycenter('face', index))]
            flag[self.face_to_dof()[index].reshape(-1).long()] = True
        elif method == 'interp':
            dofs = self.face_to_dof()[index].reshape(-1).long()
            if callable(threshold):
                dofs = dofs[threshold(self.interpolation_points()[dofs])]
            flag[dofs] = True
        else:
            raise ValueError(f"Unknown method: {method}")
        return flag

    def boundary_interpolate(self, gd, uh=None, *, threshold=None, method=None):
        """uh[bd] = gd(interpolation points on the boundary)
        (lagrange_fe_space.py:109-142).  Returns (uh, isDDof)."""
        if uh is None:
            uh = torch.zeros(self.number_of_global_dofs(), dtype=self.ftype, device=self.device)
        arr = uh.array if isinstance(uh, Function) else uh
        if threshold is None:
            isDDof, idx, pts = self._boundary()
            if torch.is_tensor(gd):
                arr[idx] = gd[idx].to(arr.dtype)
            elif callable(gd):
                val = torch.as_tensor(gd(pts), dtype=arr.dtype, device=arr.device)
                if arr.ndim == 2 and val.ndim == 1:      # one value per boundary DOF, every column
                    val = val[:, None]
                arr[idx] = val
            elif isinstance(gd, (int, float)):
                arr[idx] = gd
            else:
                raise TypeError("gd must be a tensor or a callable function")
            return (uh if isinstance(uh, Function) else Function(self, arr)), isDDof.clone()
        isDDof = self.is_boundary_dof(threshold=threshold, method='interp')
        if torch.is_tensor(gd):
            arr[isDDof] = gd[isDDof].to(arr.dtype)
        elif callable(gd):
            val = gd(self.interpolation_points()[isDDof])
            arr[isDDof] = torch.as_tensor(val, dtype=arr.dtype, device=arr.device)
        elif isinstance(gd, (int, float)):
            arr[isDDof] = gd
        else:
            raise TypeError("gd must be a tensor or a callable function")
        return (uh if isinstance(uh, Function) else Function(self, arr)), isDDof

    set_dirichlet_bc = boundary_interpolate

    def basis(self, bc, index=_S):
        phi = torch.as_tensor(shape_function(_as_points_array(bc), self.p), dtype=self.ftype, device=self.device)
        return phi[None, ...]

    def grad_basis(self, bc, index=_S, variable='x'):
        return self.mesh.grad_shape_function(bc, self.p, index=index, variables=variable)

    def value(self, uh, bc, index=_S):
        phi = self.basis(bc)[0]                                    # (NQ, ldof)
        c2d = self.cell_to_dof(index).long()
        return torch.einsum('ql, cl -> cq', phi.to(uh.dtype), uh[c2d])

    def grad_value(self, uh, bc, index=_S):
        """(NC, NQ, GD) gradient of uh at barycentric points
        (lagrange_fe_space.py:226-235)."""
        gphi = self.grad_basis(bc, index=index, variable='x')        # (NC, NQ, ldof, GD)
        c2d = self.cell_to_dof(index).long()
        return torch.einsum('cqlm, cl -> cqm', gphi.to(uh.dtype), uh[c2d])

    def function(self, array=None, dtype=None):
        return Function(self, array, dtype)

    def array(self, dim=None, dtype=None):
        n = self.number_of_global_dofs()
        shape = (n,) if dim is None else ((n, dim) if isinstance(dim, int) else (n,) + tuple(dim))
        return torch.zeros(shape, dtype=dtype or self.ftype, device=self.device)

    def interpolate(self, u):
        return Function(self, torch.as_tensor(u(self.interpolation_points()), dtype=self.ftype, device=self.device))

    # -- reference-element tables for the kernels -------------------------------
    def tables(self, q):
        """(qw, phi, gphi_u) on the device, FP64, for quadrature index q."""
        key = int(q)
        if key not in self._tables:
            qf = self.mesh.quadrature_formula(q)
            bcs, ws = qf.get_quadrature_points_and_weights()
            bc = _as_points_array(bcs)
            dev = self.device
            qw = torch.as_tensor(_as_points_array(ws), dtype=torch.float64, device=dev).contiguous()
            phi = torch.as_tensor(shape_function(bc, self.p), dtype=torch.float64, device=dev).contiguous()
            gphi = torch.as_tensor(grad_shape_function(bc, self.p), dtype=torch.float64, device=dev).contiguous()
            self._tables[key] = (bcs, qw, phi, gphi)
        return self._tables[key]

    def bcs_device(self, q):
        """Barycentric points of rule q, (NQ, TD+1) FP64 on the device."""
        key = ('bcs', int(q))
        if key not in self._tables:
            bcs = self.tables(q)[0]
            self._tables[key] = torch.as_tensor(_as_points_array(bcs), dtype=torch.float64,
                                                device=self.device).contiguous()
        return self._tables[key]

    def multi_index_matrix(self):
        return multi_index_matrix(self.p, self.TD)
