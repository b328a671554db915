"""Dirichlet boundary condition (fealpy/fem/dirichlet_bc.py:13-234)."""
import threading

import torch

from .. import _native as nat
from ..sparse.tensors import COOTensor, CSRTensor


_PATTERN_LOCK = threading.Lock()


class DirichletBC:
    def __init__(self, space, gd=None, *, threshold=None, method=None):
        if isinstance(space, tuple):
            raise NotImplementedError("tuples of spaces are not provided")
        self.space = space
        self.gd = gd
        self.threshold = threshold
        self.method = method
        self.bctype = 'Dirichlet'
        self.gdof = space.number_of_global_dofs()
        if torch.is_tensor(threshold):
            self.is_boundary_dof = threshold
            self._bd_u8 = threshold.to(torch.uint8).contiguous()
        elif threshold is None and hasattr(space, '_boundary'):
            # boundary flags depend only on the mesh: shared by every BC of the space
            flag = space._boundary()[0]
            self.is_boundary_dof = flag
            if getattr(space, '_bd_u8', None) is None:
                space._bd_u8 = flag.to(torch.uint8).contiguous()
            self._bd_u8 = space._bd_u8
        else:
            self.is_boundary_dof = space.is_boundary_dof(threshold=threshold, method=method)
            self._bd_u8 = self.is_boundary_dof.to(torch.uint8).contiguous()

    @property
    def boundary_dof_index(self):
        return torch.nonzero(self.is_boundary_dof)[:, 0]

    def check_matrix(self, matrix):
        if not isinstance(matrix, (COOTensor, CSRTensor)):
            raise ValueError('The type of matrix must be COOTensor or CSRTensor.')
        if len(matrix.shape) != 2:
            raise ValueError('The matrix must be a 2-D sparse COO matrix.')
        if matrix.shape[0] != matrix.shape[1]:
            raise ValueError('The matrix must be a square matrix.')
        if matrix.shape[0] != self.gdof:
            raise ValueError('The matrix size must match the gdof of the space.')
        return matrix

    def check_vector(self, vector):
        if not torch.is_tensor(vector):
            raise ValueError('The type of vector must be a tensor.')
        if vector.ndim not in (1, 2):
            raise ValueError('The vector must be 1-D or 2-D.')
        if vector.shape[0] != self.gdof:
            raise ValueError('The vector size must match the gdof of the space.')
        return vector

    def apply(self, A, f, uh=None, gd=None, *, check=True, keep_pattern=False):
        """-> (A with boundary rows/columns removed and a unit boundary
        diagonal, f - A g with f[bd] = g[bd]); one pass over A
        (fa_dirichlet_apply), the results of apply_vector + apply_matrix.

        keep_pattern=True (not in the reference): the pattern-preserving
        mode.  The returned CSR matrix keeps A's row pointer and column
        tensors, with the unit row on boundary rows and 0 on boundary columns
        (fa_dirichlet_apply_keep); same solution and bitwise the same right-
        hand side as the filtered form, and a later boundary condition on the
        same pattern reuses everything built on it (multigrid symbolic
        products, SELL index)."""
        A0 = self.check_matrix(A) if check else A
        f = self.check_vector(f) if check else f
        if f.ndim != 1:
            if keep_pattern:
                raise ValueError("keep_pattern takes a 1-D right-hand side")
            return self.apply_matrix(A0, check=False), self.apply_vector(f, A0, uh, gd, check=False)
        gd = self.gd if gd is None else gd
        if gd is None:
            raise RuntimeError("The boundary condition is None.")
        was_coo = isinstance(A0, COOTensor)
        Ac = self._csr(A0)
        g = self._boundary_values(f, uh, gd, Ac.dtype)
        f = f.to(Ac.dtype).contiguous()
        if keep_pattern:
            vals = Ac.values.contiguous()
            new_val = torch.empty_like(vals)
            out = torch.empty_like(f)
            nat.check(nat.lib().fa_dirichlet_apply_keep(nat.dtype_code(Ac.dtype), Ac.shape[0], nat.ptr(Ac.crow),
                                                        nat.ptr(Ac.col), nat.ptr(vals), nat.ptr(self._bd_u8),
                                                        nat.ptr(new_val), nat.ptr(g), nat.ptr(f), nat.ptr(out),
                                                        nat.stream()), "fa_dirichlet_apply_keep")
            A2 = CSRTensor(Ac.crow, Ac.col, new_val, Ac.shape)
            return (A2.tocoo() if was_coo else A2), out
        new_ptr, new_idx = self._filtered_pattern(Ac)
        new_val = torch.empty(new_idx.shape[0], dtype=Ac.dtype, device=Ac.device)
        out = torch.empty_like(f)
        nat.check(nat.lib().fa_dirichlet_apply(nat.dtype_code(Ac.dtype), Ac.shape[0], nat.ptr(Ac.crow),
                                               nat.ptr(Ac.col), nat.ptr(Ac.values), nat.ptr(self._bd_u8),
                                               nat.ptr(new_ptr), nat.ptr(new_val), nat.ptr(g), nat.ptr(f),
                                               nat.ptr(out), nat.stream()), "fa_dirichlet_apply")
        A2 = CSRTensor(new_ptr, new_idx, new_val, Ac.shape)
        return (A2.tocoo() if was_coo else A2), out

    def _boundary_values(self, f, uh, gd, dtype):
        if uh is None:
            uh = torch.zeros(f.shape, dtype=dtype, device=f.device)
        uh, _ = self.space.boundary_interpolate(gd=gd, uh=uh, threshold=self.threshold, method=self.method)
        g = uh.array if hasattr(uh, 'array') else uh
        if g.shape != f.shape:
            raise ValueError(f"uh of shape {tuple(g.shape)} does not match f of shape {tuple(f.shape)}")
        return g.to(dtype).contiguous()

    def _csr(self, A):
        return A.coalesce().tocsr() if isinstance(A, COOTensor) else A

    def apply_matrix(self, matrix, *, check=True, dtype=None):
        """The filtered matrix (dirichlet_bc.py:123-147).  ``dtype``
        (default: the matrix's) may be float32 for a float64 matrix: the
        filtered operator of a mixed-precision solve, each value rounded once
        in the same pass."""
        A = self.check_matrix(matrix) if check else matrix
        was_coo = isinstance(A, COOTensor)
        A = self._csr(A)
        L = nat.lib()
        n = A.shape[0]
        dev = A.device
        dt = A.dtype if dtype is None else dtype
        new_ptr, new_idx = self._filtered_pattern(A)
        new_val = torch.empty(new_idx.shape[0], dtype=dt, device=dev)
        nat.check(L.fa_dirichlet_fill(nat.dtype_code(A.dtype), nat.dtype_code(dt), n, nat.ptr(A.crow),
                                      nat.ptr(A.col), nat.ptr(A.values), nat.ptr(self._bd_u8), nat.ptr(new_ptr),
                                      None, nat.ptr(new_val), nat.stream()), "fa_dirichlet_fill")
        out = CSRTensor(new_ptr, new_idx, new_val, A.shape)
        return out.tocoo() if was_coo else out

    def _filtered_pattern(self, A):
        """(indptr, indices) of the filtered pattern.  It depends only on A's
        pattern and the boundary flags, so it is built once, completed on the
        building stream, and kept on the space for later assemblies (also from
        other threads and streams)."""
        with _PATTERN_LOCK:
            cache = self.space.__dict__.setdefault('_bc_patterns', [])
            for (crow, col, bd), val in cache:
                if crow is A.crow and col is A.col and (bd is self._bd_u8 or torch.equal(bd, self._bd_u8)):
                    return val
            L = nat.lib()
            n = A.shape[0]
            dev = A.device
            new_ptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
            ws = nat.workspace(L.fa_dirichlet_work_bytes(n), dev)
            nnz = nat.ctypes.c_int64(0)
            nat.check(L.fa_dirichlet_count(n, nat.ptr(A.crow), nat.ptr(A.col), nat.ptr(self._bd_u8),
                                           nat.ptr(new_ptr), nat.ctypes.byref(nnz), nat.ptr(ws), ws.numel(),
                                           nat.stream()), "fa_dirichlet_count")
            new_idx = torch.empty(int(nnz.value), dtype=torch.int32, device=dev)
            scratch = torch.empty(int(nnz.value), dtype=A.dtype, device=dev)
            nat.check(L.fa_dirichlet_fill(nat.dtype_code(A.dtype), nat.dtype_code(A.dtype), n, nat.ptr(A.crow),
                                          nat.ptr(A.col), nat.ptr(A.values), nat.ptr(self._bd_u8), nat.ptr(new_ptr),
                                          nat.ptr(new_idx), nat.ptr(scratch), nat.stream()), "fa_dirichlet_fill")
            torch.cuda.current_stream(dev).synchronize()
            del cache[:-3]                      # keep a few patterns
            cache.append(((A.crow, A.col, self._bd_u8), (new_ptr, new_idx)))
            return new_ptr, new_idx

    def apply_vector(self, vector, matrix, uh=None, gd=None, *, check=True):
        A = self.check_matrix(matrix) if check else matrix
        f = self.check_vector(vector) if check else vector
        gd = self.gd if gd is None else gd
        if gd is None:
            raise RuntimeError("The boundary condition is None.")
        A = self._csr(A)
        g = self._boundary_values(f, uh, gd, A.dtype)
        f = f.to(A.dtype)
        if f.ndim == 2:             # one right-hand side per column: f - A uh column by column
            cols = [self._rhs(A, f[:, j].contiguous(), g[:, j].contiguous()) for j in range(f.shape[1])]
            return torch.stack(cols, dim=1) if cols else torch.empty_like(f)
        return self._rhs(A, f.contiguous(), g)

    def _rhs(self, A, f, g):
        """f - A g with f[bd] = g[bd] for one contiguous column (fa_dirichlet_rhs)."""
        out = torch.empty_like(f)
        nat.check(nat.lib().fa_dirichlet_rhs(nat.dtype_code(A.dtype), A.shape[0], nat.ptr(A.crow), nat.ptr(A.col),
                                             nat.ptr(A.values), nat.ptr(self._bd_u8), nat.ptr(g), nat.ptr(f),
                                             nat.ptr(out), nat.stream()), "fa_dirichlet_rhs")
        return out
