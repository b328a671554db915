"""Integrators of the reference's plugin protocol (fealpy/fem/integrator.py:67-207).

Every integrator offers ``assembly(space, indices=None)`` -> element tensor and
``to_global_dof(space, indices=None)`` -> (NC, ldof), so user integrators that
follow the same protocol plug into ``BilinearForm``/``LinearForm`` unchanged.
The element tensors of the built-in integrators are computed by the HIP
element kernels (``fa_element_matrices`` / ``fa_element_vectors``).
"""
import torch

from .. import _native as nat
from ..expr import Affine, Sep

_S = slice(None)


class Integrator:
    _region = None
    etype = 'cell'

    def __init__(self):
        self._cache = {}
        self._keep_data = False

    def keep_data(self, status_on=True, /):
        self._keep_data = status_on
        if not status_on:
            self._cache.clear()
        return self

    def clear(self):
        self._cache.clear()

    def set_region(self, region, /):
        self._region = region
        self.clear()
        return self

    def get_region(self):
        return self._region

    def entity_selection(self, indices=None, *, mesh=None):
        if self._region is None:
            return _S if indices is None else indices
        region = self._region(mesh) if callable(self._region) else self._region
        if indices is None:
            return region
        if region.dtype == torch.bool:
            return torch.nonzero(region)[:, 0][indices]
        return region[indices]

    def size(self, mesh, /):
        if self._region is None:
            return mesh.count(self.etype)
        region = self._region(mesh) if callable(self._region) else self._region
        return int(region.sum()) if region.dtype == torch.bool else region.shape[0]

    def __call__(self, *args, **kwargs):
        return self.assembly(*args, **kwargs)

    def __add__(self, other):
        if isinstance(other, Integrator):
            return GroupIntegrator(self, other)
        return NotImplemented

    __iadd__ = __add__

    def __repr__(self):
        return self.__class__.__name__


class GroupIntegrator(Integrator):
    """Sum of integrators sharing one cell->DOF map (integrator.py:295-378)."""

    def __init__(self, *ints, region=None):
        super().__init__()
        self.ints = []
        if not ints:
            raise ValueError("No integrators provided.")
        for it in ints:
            if isinstance(it, GroupIntegrator):
                self.ints.extend(it.ints)
            elif isinstance(it, Integrator):
                self.ints.append(it)
            else:
                raise TypeError(f"Unsupported type {it.__class__.__name__} found in the inputs.")
        if region is not None:
            self.set_region(region)

    def __iter__(self):
        return iter(self.ints)

    def __len__(self):
        return len(self.ints)

    def __getitem__(self, i):
        return self.ints[i]

    def __iadd__(self, other):
        if isinstance(other, GroupIntegrator):
            self.ints.extend(other.ints)
        elif isinstance(other, Integrator):
            self.ints.append(other)
        else:
            return NotImplemented
        return self

    def set_region(self, region, /):
        for it in self.ints:
            it.set_region(region)
        return super().set_region(region)

    def to_global_dof(self, space, /, indices=None):
        return self.ints[0].to_global_dof(space, indices=indices)

    def assembly(self, space, /, indices=None, dtype=None):
        fused = _fused_scalar_group(self.ints)
        if fused is not None:
            return _scalar_elements(space, fused, indices=indices, selection=self.ints[0].entity_selection(indices),
                                    dtype=dtype)
        ct = _call_assembly(self.ints[0], space, indices, dtype)
        for it in self.ints[1:]:
            new = _call_assembly(it, space, indices, dtype)
            if new.shape[-3:] != ct.shape[-3:]:
                raise RuntimeError(f"The output of the integrator {it.__class__.__name__} has an incompatible "
                                   f"shape {tuple(new.shape)} with the previous {tuple(ct.shape)}.")
            ct = _axpby_sum(ct, new)
        return ct


def _call_assembly(it, space, indices, dtype):
    try:
        return it.assembly(space, indices=indices, dtype=dtype)
    except TypeError:
        out = it.assembly(space) if indices is None else it.assembly(space, indices=indices)
        return out if dtype is None else out.to(dtype)


def _axpby_sum(a, b):
    """a + b elementwise on the HIP vector kernel."""
    b = b.to(a.dtype).contiguous()
    out = a.contiguous().clone()
    nat.check(nat.lib().fa_axpby(nat.dtype_code(out.dtype), out.numel(), 1.0, nat.ptr(b), 1.0, nat.ptr(out),
                                 nat.stream()), "fa_axpby")
    return out


# ---- coefficients (fealpy/utils/utils.py:21-53) ---------------------------------
def coef_spec(coef, space, q, selection=_S, allow_sep=False):
    """-> (kind, scale, tensor-or-None) for the element kernels.  A separable
    expression (fealpy_amd.expr.Sep) goes to kernels that evaluate it
    in registers (`allow_sep`, all cells); elsewhere it is evaluated with
    torch at the quadrature points like any cartesian callable."""
    if coef is None:
        return nat.COEF_CONST, 1.0, None
    if isinstance(coef, (int, float)):
        return nat.COEF_CONST, float(coef), None
    mesh = space.mesh
    if isinstance(coef, Affine):
        coef = Sep.of(coef)
    if isinstance(coef, Sep):
        if coef.max_dim() >= mesh.GD:
            raise ValueError(f"the expression uses coordinate {coef.max_dim()} of a {mesh.GD}-D mesh")
        if allow_sep and isinstance(selection, slice) and selection == _S:
            return nat.COEF_SEP, 1.0, coef.device_table(space.bcs_device(q), mesh.number_of_cells())
    if callable(coef) and not torch.is_tensor(coef):
        bcs = space.tables(q)[0]
        if getattr(coef, 'coordtype', 'barycentric') == 'barycentric':
            val = coef(bcs.to(space.device), index=selection)
        elif isinstance(coef, Sep):
            val = coef(mesh.bc_to_point(bcs, index=selection), index=None if selection is _S else selection)
        else:
            val = coef(mesh.bc_to_point(bcs, index=selection))
        coef = val
        if isinstance(coef, (int, float)):
            return nat.COEF_CONST, float(coef), None
    t = torch.as_tensor(coef, device=space.device)
    if t.numel() == 1:
        return nat.COEF_CONST, float(t.reshape(-1)[0].item()), None
    t = t.to(torch.float64).contiguous()
    nc = mesh.cell[selection].shape[0]
    nq = space.tables(q)[1].shape[0]
    if t.ndim == 1 and t.shape[0] == nc:
        return nat.COEF_CELL, 1.0, t
    if t.ndim == 2 and tuple(t.shape) == (nc, nq):
        return nat.COEF_CELLQ, 1.0, t
    if t.ndim == 2 and t.shape[0] == nc and t.shape[1] == 1:
        return nat.COEF_CELL, 1.0, t.reshape(-1).contiguous()
    raise ValueError(f"coefficient of shape {tuple(t.shape)} is not supported (expected (NC,) or (NC, NQ))")


def _fused_scalar_group(ints):
    """If every integrator is a scalar diffusion/mass term with the same q,
    return [(op, integrator), ...] for a single fused element-kernel launch."""
    ops = []
    for it in ints:
        if type(it) is ScalarDiffusionIntegrator:
            ops.append(('diff', it))
        elif type(it) is ScalarMassIntegrator:
            ops.append(('mass', it))
        else:
            return None
    kinds = [o for o, _ in ops]
    if len(set(kinds)) != len(kinds) or len({it.q for _, it in ops}) != 1:
        return None
    return ops


MFMA_METHODS = {'mfma_bf16': 'bf16', 'mfma_fp32': 'fp32'}


def _scalar_elements(space, ops, *, indices=None, selection=_S, dtype=None):
    L = nat.lib()
    mesh = space.mesh
    q = ops[0][1].q if ops[0][1].q is not None else space.p + 3
    methods = {getattr(it, 'method', None) for _, it in ops}
    if len(methods) == 1 and next(iter(methods)) in MFMA_METHODS:
        if not (isinstance(selection, slice) and selection == _S):
            raise NotImplementedError("the MFMA element path assembles all cells")
        from .highorder import element_matrices_mfma
        spec = {'diff': None, 'mass': None}
        for op, it in ops:
            spec[op] = coef_spec(it.coef, space, q)
        elem = element_matrices_mfma(space, spec['diff'], spec['mass'], q, MFMA_METHODS[next(iter(methods))])
        # the MFMA tensor is FP32; like every other path the default value type is FP64
        return elem if dtype == torch.float32 else elem.to(torch.float64 if dtype is None else dtype)
    bcs, qw, phi, gphi = space.tables(q)
    cell = mesh.cell[selection].to(torch.int32).contiguous()
    nc = cell.shape[0]
    ldof = space.number_of_local_dofs()
    spec = {'diff': (nat.COEF_NONE, 0.0, None), 'mass': (nat.COEF_NONE, 0.0, None)}
    for op, it in ops:
        spec[op] = coef_spec(it.coef, space, q, selection, allow_sep=True)
    dtype = torch.float64 if dtype is None else dtype
    elem = torch.empty((nc, ldof, ldof), dtype=dtype, device=space.device)
    (dk, ds, dc), (mk, ms, mc) = spec['diff'], spec['mass']
    nat.check(L.fa_element_matrices(nat.dtype_code(dtype), space.TD, ldof, qw.shape[0], nat.ptr(mesh.node),
                                    nat.ptr(cell), nc, nat.ptr(qw), nat.ptr(phi), nat.ptr(gphi), dk, ds,
                                    nat.ptr(dc), mk, ms, nat.ptr(mc), nat.ptr(elem), nat.stream()),
              "fa_element_matrices")
    return elem


class _ScalarCellInt(Integrator):
    def to_global_dof(self, space, /, indices=None):
        return space.cell_to_dof(index=self.entity_selection(indices))

    def _check(self):
        if self.batched:
            raise NotImplementedError("batched coefficients are served by fealpy_amd.serving, not here")


class ScalarDiffusionIntegrator(_ScalarCellInt):
    """grad(u).grad(v) with coefficient (fealpy/fem/scalar_diffusion_integrator.py:14-63)."""

    def __init__(self, coef=None, q=None, *, region=None, batched=False, method=None):
        super().__init__()
        self.coef, self.q, self.batched, self.method = coef, q, batched, method
        self.set_region(region)

    def assembly(self, space, /, indices=None, dtype=None):
        self._check()
        return _scalar_elements(space, [('diff', self)], indices=indices, selection=self.entity_selection(indices),
                                dtype=dtype)


class ScalarMassIntegrator(_ScalarCellInt):
    """u v with coefficient (fealpy/fem/scalar_mass_integrator.py:13-54)."""

    def __init__(self, coef=None, q=None, *, index=_S, region=None, batched=False, method=None):
        super().__init__()
        self.coef, self.q, self.batched, self.method = coef, q, batched, method
        if region is None and not (isinstance(index, slice) and index == _S):
            region = index
        self.set_region(region)

    def assembly(self, space, /, indices=None, dtype=None):
        self._check()
        return _scalar_elements(space, [('mass', self)], indices=indices, selection=self.entity_selection(indices),
                                dtype=dtype)


class ScalarSourceIntegrator(_ScalarCellInt):
    """f v (fealpy/fem/scalar_source_integrator.py:13-58)."""

    def __init__(self, source=None, q=None, *, region=None, batched=False, method=None):
        super().__init__()
        self.source, self.q, self.batched, self.method = source, q, batched, method
        self.set_region(region)

    def assembly(self, space, /, indices=None, dtype=None):
        self._check()
        L = nat.lib()
        mesh = space.mesh
        sel = self.entity_selection(indices)
        q = self.q if self.q is not None else space.p + 3
        bcs, qw, phi, _ = space.tables(q)
        cell = mesh.cell[sel].to(torch.int32).contiguous()
        nc = cell.shape[0]
        ldof = space.number_of_local_dofs()
        kind, scale, src = coef_spec(self.source, space, q, sel, allow_sep=True)
        if self.source is None:
            kind, scale, src = nat.COEF_CONST, 1.0, None
        dtype = torch.float64 if dtype is None else dtype
        elem = torch.empty((nc, ldof), dtype=dtype, device=space.device)
        nat.check(L.fa_element_vectors(nat.dtype_code(dtype), space.TD, ldof, qw.shape[0], nat.ptr(mesh.node),
                                       nat.ptr(cell), nc, nat.ptr(qw), nat.ptr(phi), kind, scale, nat.ptr(src),
                                       nat.ptr(elem), nat.stream()), "fa_element_vectors")
        return elem


# ---- fused P1 path (element rows computed inside the CSR-row kernel) ------------
def p1_fusable_ops(it):
    """[(op, integrator)] when `it` is a scalar diffusion/mass integrator (or a
    group of them) over all cells, so its P1 matrix can be assembled by
    fa_assemble_p1_matrix; else None."""
    ints = it.ints if isinstance(it, GroupIntegrator) else [it]
    if any(i.get_region() is not None or getattr(i, 'batched', False) for i in ints):
        return None
    if isinstance(it, GroupIntegrator):
        return _fused_scalar_group(ints)
    if type(it) is ScalarDiffusionIntegrator:
        return [('diff', it)]
    if type(it) is ScalarMassIntegrator:
        return [('mass', it)]
    return None


# Meshes of at least this many cells assemble P1 in two passes (per cell, then
# per row over the incidence slots); smaller ones use the single fused kernel,
# whose one launch and no slot buffer cost less when the mesh is small.
TWO_PASS_MIN_CELLS = 65536
# P1 matrices and load vectors of TetrahedronMesh.from_box meshes take the
# structured kernels (fa_p1_box_matrix, fa_p1_box_vector); False routes them
# to the fused / two-pass kernels like any other mesh
BOX_ASSEMBLY = True


def _box_grid(space, pat):
    """(nx, ny, nz) when the P1 matrix can take the structured one-pass kernel:
    a TetrahedronMesh.from_box mesh (no threshold) and the space's own pattern."""
    mesh = space.mesh
    fb = mesh.meshdata.get('from_box')
    if not BOX_ASSEMBLY or fb is None or fb[0] != 'tet' or space.p != 1:
        return None
    nx, ny, nz = fb[2], fb[3], fb[4]
    if pat.gdof != (nx + 1) * (ny + 1) * (nz + 1) or pat.nc != 6 * nx * ny * nz or pat.nc != mesh.number_of_cells():
        return None
    return nx, ny, nz


def p1_matrix_values(space, ops, pat, dtype):
    """CSR values of one fusable group on the space's pattern (P1)."""
    L = nat.lib()
    mesh = space.mesh
    q = ops[0][1].q if ops[0][1].q is not None else space.p + 3
    _, qw, phi, _ = space.tables(q)
    cell = space.cell_to_dof()
    spec = {'diff': (nat.COEF_NONE, 0.0, None), 'mass': (nat.COEF_NONE, 0.0, None)}
    for op, it in ops:
        spec[op] = coef_spec(it.coef, space, q, allow_sep=True)
    (dk, ds, dc), (mk, ms, mc) = spec['diff'], spec['mass']
    values = torch.empty(pat.nnz, dtype=dtype, device=space.device)
    grid = _box_grid(space, pat)
    if grid is not None:         # from_box tetrahedra: one pass, cells found from each node's grid position
        nat.check(L.fa_p1_box_matrix(nat.dtype_code(dtype), *grid, nat.ptr(mesh.node_rows16()),
                                     nat.ptr(mesh.box_axes()), qw.shape[0],
                                     nat.ptr(qw), nat.ptr(phi), dk, ds, nat.ptr(dc), mk, ms, nat.ptr(mc),
                                     nat.ptr(pat.indptr), pat.nnz, nat.ptr(values), nat.stream()), "fa_p1_box_matrix")
        return values
    maps = pat.incidence_maps() if pat.nc >= TWO_PASS_MIN_CELLS else None
    if maps is not None:         # pass 1 per cell into incidence slots, pass 2 streams each row's slots
        # an FP32 matrix keeps its element rows in FP32 (half the slot traffic)
        rows = torch.empty(pat.nc * pat.ldof * pat.ldof, dtype=dtype, device=space.device)
        rcode = nat.dtype_code(dtype)
        nat.check(L.fa_p1_incidence_rows(rcode, space.TD, pat.nc, qw.shape[0], nat.ptr(mesh.node_rows16()),
                                         nat.ptr(cell), nat.ptr(qw), nat.ptr(phi), dk, ds, nat.ptr(dc), mk, ms,
                                         nat.ptr(mc), nat.ptr(maps[0]), nat.ptr(rows), nat.stream()),
                  "fa_p1_incidence_rows")
        nat.check(L.fa_sum_incidence_rows(nat.dtype_code(dtype), rcode, space.TD, pat.gdof, nat.ptr(pat.adj_ptr),
                                          nat.ptr(pat.indptr), nat.ptr(maps[1]), nat.ptr(rows), pat.max_row,
                                          nat.ptr(values), nat.stream()), "fa_sum_incidence_rows")
        return values
    nat.check(L.fa_assemble_p1_matrix(nat.dtype_code(dtype), space.TD, pat.gdof, qw.shape[0], nat.ptr(mesh.node),
                                      nat.ptr(cell), nat.ptr(qw), nat.ptr(phi), dk, ds, nat.ptr(dc), mk, ms,
                                      nat.ptr(mc), nat.ptr(pat.adj_ptr), nat.ptr(pat.adj), nat.ptr(pat.indptr),
                                      nat.ptr(pat.indices), max(pat.max_row, 1), nat.ptr(values), nat.stream()),
              "fa_assemble_p1_matrix")
    return values


def p1_source_fusable(it):
    return type(it) is ScalarSourceIntegrator and it.get_region() is None and not it.batched


def p1_vector_values(space, it, pat, dtype):
    L = nat.lib()
    mesh = space.mesh
    q = it.q if it.q is not None else space.p + 3
    _, qw, phi, _ = space.tables(q)
    kind, scale, src = coef_spec(it.source, space, q, allow_sep=True)
    if it.source is None:
        kind, scale, src = nat.COEF_CONST, 1.0, None
    out = torch.empty(pat.gdof, dtype=dtype, device=space.device)
    grid = _box_grid(space, pat)
    if grid is not None:         # from_box tetrahedra: cell-order element vectors, then a gather per node by grid position
        work = nat.workspace(L.fa_p1_box_vector_work_bytes(*grid, qw.shape[0]), space.device)
        nat.check(L.fa_p1_box_vector(nat.dtype_code(dtype), *grid, nat.ptr(mesh.node_rows16()), qw.shape[0],
                                     nat.ptr(qw), nat.ptr(phi), kind, scale, nat.ptr(src), nat.ptr(out),
                                     nat.ptr(work), work.numel(), nat.stream()), "fa_p1_box_vector")
        return out
    maps = pat.incidence_maps() if pat.nc >= TWO_PASS_MIN_CELLS else None
    if maps is not None:         # one value per incidence slot, then an ordered sum per row
        vals = torch.empty(pat.nc * pat.ldof, dtype=torch.float64, device=space.device)
        nat.check(L.fa_p1_incidence_vectors(space.TD, pat.nc, qw.shape[0], nat.ptr(mesh.node_rows16()),
                                            nat.ptr(space.cell_to_dof()), nat.ptr(qw), nat.ptr(phi), kind, scale,
                                            nat.ptr(src), nat.ptr(maps[0]), nat.ptr(vals), nat.stream()),
                  "fa_p1_incidence_vectors")
        nat.check(L.fa_sum_incidence_values(nat.dtype_code(dtype), pat.gdof, nat.ptr(pat.adj_ptr), nat.ptr(vals),
                                            nat.ptr(out), nat.stream()), "fa_sum_incidence_values")
        return out
    nat.check(L.fa_assemble_p1_vector(nat.dtype_code(dtype), space.TD, pat.gdof, qw.shape[0], nat.ptr(mesh.node),
                                      nat.ptr(space.cell_to_dof()), nat.ptr(qw), nat.ptr(phi), kind, scale,
                                      nat.ptr(src), nat.ptr(pat.adj_ptr), nat.ptr(pat.adj), nat.ptr(out),
                                      nat.stream()), "fa_assemble_p1_vector")
    return out
