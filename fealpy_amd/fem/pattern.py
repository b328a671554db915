"""Sparsity pattern of a cell->DOF map, built on the GPU and cached per map.

Holds the DOF->(cell, local) incidence (used by every value gather) and the
CSR pattern (int32 indptr/indices, columns sorted in each row).
"""
import threading

import torch

from .. import _native as nat


class Pattern:
    __slots__ = ("cell2dof", "gdof", "ldof", "nc", "adj_ptr", "adj", "max_degree", "indptr", "indices", "nnz",
                 "max_row", "_maps", "_maps_lock")

    def __init__(self, cell2dof, gdof):
        L = nat.lib()
        c2d = cell2dof.to(torch.int32).contiguous()
        nc, ldof = c2d.shape
        dev = c2d.device
        self.cell2dof, self.gdof, self.ldof, self.nc = c2d, int(gdof), int(ldof), int(nc)
        self.adj_ptr = torch.empty(gdof + 1, dtype=torch.int32, device=dev)
        self.adj = torch.empty(max(nc * ldof, 1), dtype=torch.int32, device=dev)
        ws = nat.workspace(L.fa_adjacency_work_bytes(nc, ldof, gdof), dev)
        maxdeg = nat.ctypes.c_int32(0)
        nat.check(L.fa_adjacency_build(nat.ptr(c2d), nc, ldof, gdof, nat.ptr(self.adj_ptr), nat.ptr(self.adj),
                                       nat.ctypes.byref(maxdeg), nat.ptr(ws), ws.numel(), nat.stream()),
                  "fa_adjacency_build")
        self.max_degree = int(maxdeg.value)
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
        self._maps = None
        self._maps_lock = threading.Lock()

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


def pattern_of(space, cell2dof=None):
    """Cached pattern of ``space.cell_to_dof()`` (or of an explicit map).  A
    new pattern is completed on the building stream before it is published,
    so other threads and streams may use it at once."""
    if cell2dof is None:
        if getattr(space, "_pattern", None) is None:
            with _LOCK:
                if getattr(space, "_pattern", None) is None:
                    pat = Pattern(space.cell_to_dof(), space.number_of_global_dofs())
                    torch.cuda.current_stream(pat.indptr.device).synchronize()
                    space._pattern = pat
        return space._pattern
    key = (cell2dof.data_ptr(), tuple(cell2dof.shape))
    cache = space.__dict__.setdefault("_patterns", {})
    if key not in cache:
        cache[key] = Pattern(cell2dof, space.number_of_global_dofs())
    return cache[key]
