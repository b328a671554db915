"""SELL-64 copies of CSR matrices for the solver kernels (see include/fealpy_amd.h)."""
import ctypes
import threading

import torch

from .. import _native as nat


class FaSell(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("ncols", ctypes.c_int64), ("slice_ptr", ctypes.c_void_p),
                ("cols", ctypes.c_void_p), ("vals", ctypes.c_void_p)]


class FaSell16(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("slice_ptr", ctypes.c_void_p), ("offs", ctypes.c_void_p),
                ("vals", ctypes.c_void_p)]


_INDEX_LOCK = threading.Lock()


def _build_index(crow, col, shape):
    L = nat.lib()
    n = shape[0]
    dev = crow.device
    slice_ptr = torch.empty((n + 63) // 64 + 1, dtype=torch.int32, device=dev)
    ws = nat.workspace(L.fa_sell_work_bytes(n), dev)
    cnt = ctypes.c_int64(0)
    nat.check(L.fa_sell_count(n, nat.ptr(crow), nat.ptr(slice_ptr), ctypes.byref(cnt), nat.ptr(ws), ws.numel(),
                              nat.stream()), "fa_sell_count")
    entries = int(cnt.value)
    cols = torch.zeros(max(entries, 1), dtype=torch.int32, device=dev)
    if entries > 0:
        nat.check(L.fa_sell_fill(nat.FA_F64, nat.FA_F64, n, shape[1], nat.ptr(crow), nat.ptr(col), None,
                                 nat.ptr(slice_ptr), nat.ptr(cols), None, nat.stream()), "fa_sell_fill")
    return slice_ptr, cols, entries


def _sell_index(crow, col, shape):
    """(slice_ptr, cols, entries) of a CSR pattern, built once per pattern and
    completed on the building stream before it is shared.  The index is kept
    on the pattern's column tensor, so it lives exactly as long as the
    pattern: every matrix of one pattern shares it, and a dropped pattern
    takes its index with it."""
    with _INDEX_LOCK:
        ent = getattr(col, "_fa_sell_index", None)
        if ent is not None and ent[0] is crow and ent[1] == tuple(shape):
            return ent[2]
        idx = _build_index(crow, col, shape)
        torch.cuda.current_stream(crow.device).synchronize()
        col._fa_sell_index = (crow, tuple(shape), idx)
        return idx


class SellMatrix:
    """Sliced-ELL (64-row slices) index structure of a CSR pattern plus value
    arrays in FP64 and/or FP32.  The index part is shared by every matrix of
    the same pattern; ``set_values`` refreshes this matrix's values (with an
    optional FP64->FP32 cast)."""

    def __init__(self, crow, col, shape, shared=False):
        self.crow, self.col = crow, col
        self.shape = tuple(shape)
        self.slice_ptr, self.cols, self.entries = (_sell_index if shared else _build_index)(crow, col, self.shape)
        self.vals = {}
        self.h_offs = self.h_vals = None

    def set_compressed(self, values, work):
        """Compressed FP16 D^-1 A on these slices (fa_sell16_fill).  With int16
        column offsets when every column lies within 32767 of its row; else
        the FP16 values use this matrix's int32 columns (`h_offs` None) and
        fa_smoother_setup writes them.  False only for an empty matrix."""
        if self.entries == 0:
            return False
        if self.h_vals is None:
            self.h_vals = torch.zeros(self.entries, dtype=torch.float16, device=values.device)
        band = getattr(self, "_h_band_ok", None)
        if band is False:
            return True
        checked = band is True
        if not checked:
            self.h_offs = torch.zeros(self.entries, dtype=torch.int16, device=values.device)
        st = nat.lib().fa_sell16_fill(nat.dtype_code(values.dtype), self.shape[0], nat.ptr(self.crow),
                                      nat.ptr(self.col), nat.ptr(values.contiguous()), nat.ptr(self.slice_ptr),
                                      nat.ptr(self.h_offs), nat.ptr(self.h_vals), None if checked else nat.ptr(work),
                                      0 if checked else work.numel(), nat.stream())
        if st == nat.FA_ECAPACITY:
            self.h_offs = None
            self._h_band_ok = False
            return True
        nat.check(st, "fa_sell16_fill")
        self._h_band_ok = True
        return True

    def desc16(self):
        if self.h_vals is None:
            return FaSell16(0, None, None, None)
        offs = self.h_offs.data_ptr() if self.h_offs is not None else None
        return FaSell16(self.shape[0], self.slice_ptr.data_ptr(), offs, self.h_vals.data_ptr())

    def value_buffer(self, dtype, device):
        """The SELL value array of `dtype` (allocated on first use)."""
        out = self.vals.get(dtype)
        if out is None:
            out = torch.zeros(max(self.entries, 1), dtype=dtype, device=device)
            self.vals[dtype] = out
        return out

    def set_values(self, values, dtype=None):
        """Copy CSR values (FP64 or FP32) into the SELL layout as `dtype`."""
        dtype = values.dtype if dtype is None else dtype
        out = self.value_buffer(dtype, values.device)
        if self.entries == 0 or values.numel() == 0:      # no stored entries (e.g. an empty transfer)
            return out
        nat.check(nat.lib().fa_sell_fill(nat.dtype_code(values.dtype), nat.dtype_code(dtype), self.shape[0],
                                         self.shape[1], nat.ptr(self.crow), nat.ptr(self.col), nat.ptr(values.contiguous()),
                                         nat.ptr(self.slice_ptr), None, nat.ptr(out), nat.stream()), "fa_sell_fill")
        return out

    def desc(self, dtype):
        v = self.vals[dtype]
        return FaSell(self.shape[0], self.shape[1], self.slice_ptr.data_ptr(), self.cols.data_ptr(), v.data_ptr())

    def matvec(self, x, out=None):
        y = torch.empty(self.shape[0], dtype=x.dtype, device=x.device) if out is None else out
        d = self.desc(x.dtype)
        nat.check(nat.lib().fa_sell_spmv(nat.dtype_code(x.dtype), ctypes.byref(d), nat.ptr(x.contiguous()),
                                         nat.ptr(y), nat.stream()), "fa_sell_spmv")
        return y


def sell_of(A):
    """SELL copy of a CSRTensor, cached on the tensor (index part keyed by the
    pattern arrays; values refreshed on every call)."""
    key = (A.crow.data_ptr(), A.col.data_ptr(), A.nnz)
    S = getattr(A, "_sell", None)
    if S is None or S[0] != key:
        S = (key, SellMatrix(A.crow, A.col, A.shape, shared=True))
        A._sell = S
    sm = S[1]
    sm.set_values(A.values)
    return sm
