"""BilinearForm / LinearForm (fealpy/fem/form.py, bilinear_form.py, linear_form.py).

Assembly never forms a COO list: the CSR pattern of the space's cell->DOF map
is built once on the GPU (kernel "b") and cached; each ``assembly()`` computes
the group's element tensors (kernel "a") and sums them straight into the CSR
value array in ascending-cell order (kernel "c").
"""
import torch

from .. import _native as nat
from ..sparse.tensors import COOTensor, CSRTensor
from .integrators import GroupIntegrator, Integrator, _axpby_sum, _call_assembly, coef_spec, p1_fusable_ops, \
    p1_matrix_values, p1_source_fusable, p1_vector_values
from .pattern import pattern_of


class UniformSplitter:
    """Chunks of ``chunk_size`` cells (form.py:172-188).  Kept for API parity:
    the kernels tile cells themselves, so the chunks are assembled in one pass."""

    def __init__(self, chunk_size, /):
        self.chunk_size = chunk_size


class Form:
    def __init__(self, *space, batch_size=0):
        if len(space) == 0:
            raise ValueError("No space is given.")
        if isinstance(space[0], (tuple, list)):
            space = tuple(space[0])
        if batch_size:
            raise NotImplementedError("batched forms are served by fealpy_amd.serving")
        self._spaces = space
        self.integrators = {}
        self.splitters = {}
        self._cursor = 0
        self.batch_size = batch_size
        self.sparse_shape = self._get_sparse_shape()

    def __len__(self):
        return len(self.integrators)

    @property
    def space(self):
        return self._spaces[0] if len(self._spaces) == 1 else self._spaces

    @property
    def shape(self):
        return self.sparse_shape

    def add_integrator(self, *I, region=None, splitter=None, group=None):
        if len(I) == 0:
            return self
        if len(I) == 1 and isinstance(I[0], (tuple, list)):
            I = tuple(I[0])
        if len(I) == 1:
            I = I[0]
            if region is not None:
                I.set_region(region)
        else:
            I = GroupIntegrator(*I, region=region)
        if isinstance(splitter, int):
            splitter = UniformSplitter(splitter)
        group = f'_group_{self._cursor}' if group is None else group
        self._cursor += 1
        if group in self.integrators:
            self.integrators[group] += I
        else:
            self.integrators[group] = I
        self.splitters[group] = splitter
        return self

    def __lshift__(self, other):
        if isinstance(other, Integrator):
            return self.add_integrator(other)
        return NotImplemented

    def _groups(self, dtype):
        """[(element tensor, cell->DOF map)] per integrator group."""
        space = self._spaces[0]
        out = []
        for it in self.integrators.values():
            elem = _call_assembly(it, space, None, dtype)
            e2d = it.to_global_dof(space)
            if isinstance(e2d, (tuple, list)):
                e2d = e2d[0]
            out.append((elem, e2d))
        return out


class BilinearForm(Form):
    _M = None

    def _get_sparse_shape(self):
        n = self._spaces[0].number_of_global_dofs()
        m = self._spaces[1].number_of_global_dofs() if len(self._spaces) > 1 else n
        return (m, n)

    def check_space(self):
        if len(self._spaces) != 1:
            raise NotImplementedError("BilinearForm with two different spaces is not provided")

    def assembly(self, *, format='csr', dtype=None):
        """Assemble the global matrix ([batch, ]gdof, gdof).

        ``dtype`` (float64 default, or float32) selects the value type; element
        integrals are always evaluated in FP64."""
        self.check_space()
        if format not in ('csr', 'coo'):
            raise ValueError(f"Unsupported format {format}.")
        space = self._spaces[0]
        fused = self._p1_fused(space, dtype)
        if fused is not None:
            self._M = fused if format == 'csr' else fused.tocoo()
            return self._M
        full_map = space.cell_to_dof()
        total = None
        values = None
        pat = None
        for elem, e2d in self._groups(dtype):
            if e2d.shape == full_map.shape and e2d.data_ptr() == full_map.data_ptr():
                total = elem if total is None else _axpby_sum(total, elem)
            else:                                   # a region: gather it separately on its own pattern
                p_sub = pattern_of(space, e2d)
                v = p_sub.gather_matrix(elem)
                sub = CSRTensor(p_sub.indptr, p_sub.indices, v, self.sparse_shape)
                values = sub if values is None else values.add(sub)
        if total is not None:
            pat = pattern_of(space)
            M = CSRTensor(pat.indptr, pat.indices, pat.gather_matrix(total), self.sparse_shape)
            values = M if values is None else M.add(values)
        if values is None:
            raise RuntimeError("no integrator was added to the form")
        self._M = values if format == 'csr' else values.tocoo()
        return self._M

    def _p1_fused(self, space, dtype):
        """P1 scalar diffusion/mass groups over all cells: one fused kernel per
        group (fa_assemble_p1_matrix), no element buffer."""
        if space.p != 1 or getattr(space, 'TD', None) not in (2, 3) or space.mesh.GD != space.TD:
            return None
        groups = [p1_fusable_ops(it) for it in self.integrators.values()]
        if not groups or any(g is None for g in groups):
            return None
        pat = pattern_of(space)
        if pat.max_row > 128:
            return None
        dtype = torch.float64 if dtype is None else dtype
        values = None
        for ops in groups:
            v = p1_matrix_values(space, ops, pat, dtype)
            values = v if values is None else _axpby_sum(values, v)
        return CSRTensor(pat.indptr, pat.indices, values, self.sparse_shape)

    def __matmul__(self, u):
        """self @ u.  With an assembled matrix, its product; otherwise without
        assembling (bilinear_form.py:126-158): each group's element tensors
        applied cell by cell to the gathered u, then summed per DOF in
        ascending cell order; the form keeps no matrix (``_M`` stays None).
        A 2-D u holds one vector per row (batch first), as in the reference.
        Element tensors are recomputed on each product unless every
        integrator keeps its data (``keep_data()``)."""
        if self._M is not None:
            return self._M @ u
        self.check_space()
        if u.ndim == 2:
            return torch.stack([self._apply_local(u[b]) for b in range(u.shape[0])])
        if u.ndim != 1:
            raise ValueError("u must be a 1-D or 2-D tensor")
        return self._apply_local(u)

    def _local_tensors(self, dtype):
        keep = all(getattr(it, '_keep_data', False) for it in self.integrators.values())
        key = ('local', dtype)
        cached = getattr(self, '_local', None)
        if keep and cached is not None and cached[0] == key:
            return cached[1]
        space = self._spaces[0]
        full_map = space.cell_to_dof()
        groups = []
        for elem, e2d in self._groups(dtype):
            same = e2d.shape == full_map.shape and e2d.data_ptr() == full_map.data_ptr()
            pat = pattern_of(space) if same else pattern_of(space, e2d)
            groups.append((elem.contiguous(), pat))
        self._local = (key, groups) if keep else None
        return groups

    def _apply_local(self, u):
        L = nat.lib()
        space = self._spaces[0]
        gdof = space.number_of_global_dofs()
        if u.shape[0] != gdof:
            raise ValueError(f"u of length {u.shape[0]} for a space with {gdof} DOFs")
        dt = u.dtype if u.dtype in (torch.float64, torch.float32) else torch.float64
        x = u.to(dt).contiguous()
        out = None
        for elem, pat in self._local_tensors(dt):
            elem = elem.to(dt)
            v = torch.empty((pat.nc, pat.ldof), dtype=dt, device=x.device)
            nat.check(L.fa_element_apply(nat.dtype_code(dt), pat.nc, pat.ldof, nat.ptr(pat.cell2dof), nat.ptr(elem),
                                         elem.stride(0), nat.ptr(x), nat.ptr(v), nat.stream()), "fa_element_apply")
            y = pat.gather_vector(v)
            out = y if out is None else _axpby_sum(out, y)
        if out is None:
            raise RuntimeError("no integrator was added to the form")
        return out

    def mult(self, x, out=None):
        """Matrix-free product self @ x that never reads an assembled matrix
        (the reference declares ``mult`` without providing it,
        bilinear_form.py:107-117).  P1 scalar diffusion/mass groups on a
        simplex mesh recompute their element rows in the kernel
        (fa_p1_apply, FP64); any other form applies its element tensors
        (``__matmul__`` with no matrix)."""
        self.check_space()
        if x.ndim != 1:
            raise ValueError("mult takes a 1-D vector")
        groups = self._p1_groups()
        if groups is None:
            y = self._apply_local(x)
        else:
            y = None
            for ops in groups:
                v = self._p1_apply(ops, x.to(torch.float64).contiguous())
                y = v if y is None else _axpby_sum(y, v)
        if out is not None:
            out.copy_(y)
            return out
        return y

    def residual(self, b, x, is_bd, *, dtype=torch.float64):
        """(b - A x with 0 on the rows flagged by is_bd (uint8), the squared
        norm of that vector as a 1-element FP64 device tensor).  A is applied
        matrix-free in FP64; the result is stored as `dtype` (FP32 for the
        right-hand side of an FP32 inner solve).  For a P1 scalar form this is
        one kernel (fa_p1_apply, residual form); other forms go through
        ``mult``."""
        L = nat.lib()
        space = self._spaces[0]
        n = space.number_of_global_dofs()
        if b.shape != (n,) or x.shape != (n,) or is_bd.shape != (n,):
            raise ValueError("b, x and is_bd must be vectors over the space's DOFs")
        groups = self._p1_groups()
        norm2 = torch.empty(1, dtype=torch.float64, device=x.device)
        out = torch.empty(n, dtype=dtype, device=x.device)
        b = b.to(torch.float64).contiguous()
        x = x.to(torch.float64).contiguous()
        if groups is not None and len(groups) == 1:
            self._p1_apply(groups[0], x, b, is_bd, out, norm2)
            return out, norm2
        r = b.clone()                            # r = b - A x
        nat.check(L.fa_axpby(nat.FA_F64, n, -1.0, nat.ptr(self.mult(x).contiguous()), 1.0, nat.ptr(r), nat.stream()),
                  "fa_axpby")
        res = torch.zeros(n, dtype=torch.float64, device=x.device)
        nat.check(L.fa_refine_update(nat.FA_F64, n, nat.ptr(res), nat.ptr(r), nat.ptr(is_bd), nat.stream()),
                  "fa_refine_update")            # res = r off the Dirichlet rows, 0 on them
        ws = nat.workspace(L.fa_reduce_work_bytes(), x.device)
        nat.check(L.fa_dot(nat.FA_F64, n, nat.ptr(res), nat.ptr(res), nat.ptr(norm2), nat.ptr(ws), ws.numel(),
                           nat.stream()), "fa_dot")
        if dtype == torch.float64:
            return res, norm2
        nat.check(L.fa_cast(nat.FA_F64, nat.dtype_code(dtype), n, nat.ptr(res), nat.ptr(out), nat.stream()),
                  "fa_cast")
        return out, norm2

    def _p1_groups(self):
        space = self._spaces[0]
        if space.p != 1 or getattr(space, 'TD', None) not in (2, 3) or space.mesh.GD != space.TD:
            return None
        groups = [p1_fusable_ops(it) for it in self.integrators.values()]
        if not groups or any(g is None for g in groups):
            return None
        return groups

    def _p1_apply(self, ops, x, b=None, is_bd=None, out=None, norm2=None):
        L = nat.lib()
        space = self._spaces[0]
        mesh = space.mesh
        pat = pattern_of(space)
        q = ops[0][1].q if ops[0][1].q is not None else space.p + 3
        _, qw, phi, _ = space.tables(q)
        spec = {'diff': (nat.COEF_NONE, 0.0, None), 'mass': (nat.COEF_NONE, 0.0, None)}
        for op, it in ops:
            spec[op] = coef_spec(it.coef, space, q, allow_sep=True)
        (dk, ds, dc), (mk, ms, mc) = spec['diff'], spec['mass']
        y = torch.empty(pat.gdof, dtype=torch.float64, device=x.device) if out is None else out
        ws = nat.workspace(L.fa_p1_apply_work_bytes(), x.device) if b is not None else None
        nat.check(L.fa_p1_apply(nat.dtype_code(y.dtype), space.TD, pat.gdof, qw.shape[0], nat.ptr(mesh.node_rows16()),
                                nat.ptr(space.cell_to_dof()), nat.ptr(qw), nat.ptr(phi), dk, ds, nat.ptr(dc), mk, ms,
                                nat.ptr(mc), nat.ptr(pat.adj_ptr), nat.ptr(pat.adj), nat.ptr(x), nat.ptr(b),
                                nat.ptr(is_bd), nat.ptr(y), nat.ptr(norm2), nat.ptr(ws),
                                0 if ws is None else ws.numel(), nat.stream()), "fa_p1_apply")
        return y

    @property
    def T(self):
        return self


class LinearForm(Form):
    _V = None

    def _get_sparse_shape(self):
        return (self._spaces[0].number_of_global_dofs(),)

    def check_space(self):
        if len(self._spaces) != 1:
            raise ValueError("LinearForm should have only one space.")

    def assembly(self, *, format='dense', dtype=None):
        """Assemble the global vector: element vectors summed per DOF in
        ascending-cell order (linear_form.py:36-86)."""
        self.check_space()
        if format not in ('dense', 'coo'):
            raise ValueError(f"Unsupported format {format}.")
        space = self._spaces[0]
        full_map = space.cell_to_dof()
        out = None
        ints = list(self.integrators.values())
        if (space.p == 1 and getattr(space, 'TD', None) in (2, 3) and space.mesh.GD == space.TD and ints
                and all(map(p1_source_fusable, ints))):
            pat = pattern_of(space)
            dt = torch.float64 if dtype is None else dtype
            for it in ints:                         # fused P1 load vector, no element buffer
                v = p1_vector_values(space, it, pat, dt)
                out = v if out is None else _axpby_sum(out, v)
            ints = []
        for elem, e2d in (self._groups(dtype) if ints else []):
            pat = pattern_of(space) if (e2d.shape == full_map.shape and e2d.data_ptr() == full_map.data_ptr()) \
                else pattern_of(space, e2d)
            v = pat.gather_vector(elem)
            out = v if out is None else _axpby_sum(out, v)
        if out is None:
            raise RuntimeError("no integrator was added to the form")
        if format == 'coo':
            idx = torch.arange(out.shape[0], device=out.device, dtype=torch.int32)[None, :]
            self._V = COOTensor(idx, out, (out.shape[0],), is_coalesced=True)
        else:
            self._V = out
        return self._V
