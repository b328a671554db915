"""Sparsity pattern of a cell->DOF map, built on the GPU and cached per map.

Holds the DOF->(cell, local) incidence (used by every value gather) and the
CSR pattern (int32 indptr/indices, columns sorted in each row).
"""
import threading

import torch

from .. import _native as nat


def _box_counts(nx, ny, nz):
    """(nnz, longest row) of the P1 pattern of an nx x ny x nz Kuhn grid: a
    node's row holds itself and its neighbours along the 14 Kuhn edge
    directions that lie in the grid (each axis: low end, inside, high end)."""
    dirs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)
            if all(v >= 0 for v in (a, b, c)) or all(v <= 0 for v in (a, b, c))]
    nnz, longest = 0, 0
    for kx in ((0, 1), (1, nx - 1), (nx, 1)) if nx > 1 else ((0, 1), (1, 1)):
        for ky in ((0, 1), (1, ny - 1), (ny, 1)) if ny > 1 else ((0, 1), (1, 1)):
            for kz in ((0, 1), (1, nz - 1), (nz, 1)) if nz > 1 else ((0, 1), (1, 1)):
                row = sum(1 for a, b, c in dirs if 0 <= kx[0] + a <= nx and 0 <= ky[0] + b <= ny
                          and 0 <= kz[0] + c <= nz)
                nnz += row * kx[1] * ky[1] * kz[1]
                longest = max(longest, row)
    return nnz, longest


class Pattern:
    __slots__ = ("cell2dof", "gdof", "ldof", "nc", "_adj_ptr", "_adj", "_max_degree", "indptr", "indices", "nnz",
                 "max_row", "_maps", "_maps_lock", "_adj_lock")

    def __init__(self, cell2dof, gdof, box=None):
        """box: (nx, ny, nz) of a TetrahedronMesh.from_box P1 map, whose CSR
        pattern comes from the grid (fa_p1_box_pattern); the incidence lists
        are then built on first use."""
        L = nat.lib()
        c2d = cell2dof.to(torch.int32).contiguous()
        nc, ldof = c2d.shape
        dev = c2d.device
        self.cell2dof, self.gdof, self.ldof, self.nc = c2d, int(gdof), int(ldof), int(nc)
        self._adj_ptr = self._adj = self._max_degree = None
        self._adj_lock = threading.Lock()
        self._maps = None
        self._maps_lock = threading.Lock()
        if box is not None:
            nx, ny, nz = box
            self.nnz, self.max_row = _box_counts(nx, ny, nz)
            self.indptr = torch.empty(gdof + 1, dtype=torch.int32, device=dev)
            self.indices = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
            ws = nat.workspace(L.fa_p1_box_pattern_work_bytes(nx, ny, nz), dev)
            nat.check(L.fa_p1_box_pattern(nx, ny, nz, nat.ptr(self.indptr), nat.ptr(self.indices), self.nnz,
                                          nat.ptr(ws), ws.numel(), nat.stream()), "fa_p1_box_pattern")
            self.indices = self.indices[:self.nnz]
            return
        self._build_adj()
        self.indptr = torch.empty(gdof + 1, dtype=torch.int32, device=dev)
        ws = nat.workspace(L.fa_pattern_work_bytes(gdof), dev)
        nnz = nat.ctypes.c_int64(0)
        mrow = nat.ctypes.c_int32(0)
        nat.check(L.fa_pattern_count(nat.ptr(c2d), nc, ldof, gdof, nat.ptr(self.adj_ptr), nat.ptr(self.adj),
                                     self.max_degree, nat.ptr(self.indptr), nat.ctypes.byref(nnz),
                                     nat.ctypes.byref(mrow), nat.ptr(ws), ws.numel(), nat.stream()),
                  "fa_pattern_count")
        self.nnz, self.max_row = int(nnz.value), int(mrow.value)
        self.indices = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
        nat.check(L.fa_pattern_fill(nat.ptr(c2d), nc, ldof, gdof, nat.ptr(self.adj_ptr), nat.ptr(self.adj),
                                    self.max_degree, nat.ptr(self.indptr), nat.ptr(self.indices), nat.stream()),
                  "fa_pattern_fill")
        self.indices = self.indices[:self.nnz]

    def _build_adj(self):
        L = nat.lib()
        c2d, nc, ldof, gdof = self.cell2dof, self.nc, self.ldof, self.gdof
        dev = c2d.device
        adj_ptr = torch.empty(gdof + 1, dtype=torch.int32, device=dev)
        adj = torch.empty(max(nc * ldof, 1), dtype=torch.int32, device=dev)
        ws = nat.workspace(L.fa_adjacency_work_bytes(nc, ldof, gdof), dev)
        maxdeg = nat.ctypes.c_int32(0)
        nat.check(L.fa_adjacency_build(nat.ptr(c2d), nc, ldof, gdof, nat.ptr(adj_ptr), nat.ptr(adj),
                                       nat.ctypes.byref(maxdeg), nat.ptr(ws), ws.numel(), nat.stream()),
                  "fa_adjacency_build")
        self._adj_ptr, self._adj, self._max_degree = adj_ptr, adj, int(maxdeg.value)

    def _adjacency(self):
        """Incidence lists on first use (patterns built from the grid), complete
        on the building stream before any other stream or thread sees them."""
        if self._adj is None:
            with self._adj_lock:
                if self._adj is None:
                    self._build_adj()
                    torch.cuda.current_stream(self.cell2dof.device).synchronize()
        return self._adj_ptr, self._adj, self._max_degree

    @property
    def adj_ptr(self):
        return self._adjacency()[0]

    @property
    def adj(self):
        return self._adjacency()[1]

    @property
    def max_degree(self):
        return self._adjacency()[2]

    def incidence_maps(self):
        """(inc, pos) of the two-pass P1 assembly (fa_incidence_maps): the
        incidence slot of every (cell, local DOF) and, per incidence, the CSR
        row positions of the cell's DOFs packed one byte each.  Built once and
        completed on the building stream before it is shared; None for
        ldof > 4 or rows longer than 256 entries."""
        if self._maps is None:
            with self._maps_lock:
                if self._maps is None:
                    self._maps = self._build_maps()
        return self._maps or None

    def _build_maps(self):
        if self.ldof > 4 or self.max_row > 256 or self.nnz == 0 or self.nc == 0:
            return False
        L = nat.lib()
        dev = self.indptr.device
        inc = torch.empty(self.nc * self.ldof, dtype=torch.int32, device=dev)
        pos = torch.empty(self.nc * self.ldof, dtype=torch.int32, device=dev)
        ws = nat.workspace(64, dev)
        st = L.fa_incidence_maps(self.gdof, self.ldof, nat.ptr(self.cell2dof), nat.ptr(self.adj_ptr),
                                 nat.ptr(self.adj), nat.ptr(self.indptr), nat.ptr(self.indices), nat.ptr(inc),
                                 nat.ptr(pos), nat.ptr(ws), ws.numel(), nat.stream())
        if st == nat.FA_ECAPACITY:
            return False
        nat.check(st, "fa_incidence_maps")
        return (inc, pos)

    def gather_matrix(self, elem, values=None):
        """CSR values from element matrices (NC, ldof, ldof) in ascending-cell order."""
        L = nat.lib()
        if tuple(elem.shape) != (self.nc, self.ldof, self.ldof):
            raise ValueError(f"element tensor shape {tuple(elem.shape)} != {(self.nc, self.ldof, self.ldof)}")
        if not (elem.stride(2) == 1 and elem.stride(1) == self.ldof and elem.stride(0) >= self.ldof ** 2):
            elem = elem.contiguous()                      # cells may be padded apart; rows must be dense
        if values is None:
            values = torch.empty(self.nnz, dtype=elem.dtype, device=elem.device)
        nat.check(L.fa_gather_matrix(nat.dtype_code(elem.dtype), self.gdof, self.ldof, nat.ptr(self.cell2dof),
                                     nat.ptr(self.adj_ptr), nat.ptr(self.adj), nat.ptr(self.indptr),
                                     nat.ptr(self.indices), max(self.max_row, 1), nat.ptr(elem), elem.stride(0),
                                     nat.ptr(values), nat.stream()), "fa_gather_matrix")
        return values

    def gather_vector(self, elem, out=None):
        L = nat.lib()
        elem = elem.contiguous()
        if tuple(elem.shape) != (self.nc, self.ldof):
            raise ValueError(f"element vector shape {tuple(elem.shape)} != {(self.nc, self.ldof)}")
        if out is None:
            out = torch.empty(self.gdof, dtype=elem.dtype, device=elem.device)
        nat.check(L.fa_gather_vector(nat.dtype_code(elem.dtype), self.gdof, self.ldof, nat.ptr(self.adj_ptr),
                                     nat.ptr(self.adj), nat.ptr(elem), nat.ptr(out), nat.stream()),
                  "fa_gather_vector")
        return out


_LOCK = threading.Lock()
# P1 patterns of TetrahedronMesh.from_box meshes come from the grid
# (fa_p1_box_pattern); False builds them from the incidence lists like any other
BOX_PATTERN = True


def _box_grid_of(space):
    mesh = space.mesh
    fb = mesh.meshdata.get('from_box')
    if not BOX_PATTERN or fb is None or fb[0] != 'tet' or space.p != 1:
        return None
    nx, ny, nz = fb[2], fb[3], fb[4]
    if mesh.number_of_cells() != 6 * nx * ny * nz or space.number_of_global_dofs() != (nx + 1) * (ny + 1) * (nz + 1):
        return None
    return nx, ny, nz


def pattern_of(space, cell2dof=None):
    """Cached pattern of ``space.cell_to_dof()`` (or of an explicit map).  A
    new pattern is completed on the building stream before it is published,
    so other threads and streams may use it at once."""
    if cell2dof is None:
        if getattr(space, "_pattern", None) is None:
            with _LOCK:
                if getattr(space, "_pattern", None) is None:
                    pat = Pattern(space.cell_to_dof(), space.number_of_global_dofs(), box=_box_grid_of(space))
                    torch.cuda.current_stream(pat.indptr.device).synchronize()
                    space._pattern = pat
        return space._pattern
    key = (cell2dof.data_ptr(), tuple(cell2dof.shape))
    cache = space.__dict__.setdefault("_patterns", {})
    if key not in cache:
        cache[key] = Pattern(cell2dof, space.number_of_global_dofs())
    return cache[key]
