"""SELL-64 copies of CSR matrices for the solver kernels (see include/fealpy_amd.h)."""
import ctypes

import torch

from .. import _native as nat


class FaSell(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("ncols", ctypes.c_int64), ("slice_ptr", ctypes.c_void_p),
                ("cols", ctypes.c_void_p), ("vals", ctypes.c_void_p)]


class SellMatrix:
    """Sliced-ELL (64-row slices) index structure of a CSR pattern plus value
    arrays in FP64 and/or FP32.  The index part is built once per pattern;
    ``set_values`` refreshes the values (with an optional FP64->FP32 cast)."""

    def __init__(self, crow, col, shape):
        L = nat.lib()
        self.crow, self.col = crow, col
        self.shape = tuple(shape)
        n = self.shape[0]
        dev = crow.device
        self.slice_ptr = torch.empty((n + 63) // 64 + 1, dtype=torch.int32, device=dev)
        ws = nat.workspace(L.fa_sell_work_bytes(n), dev)
        ent = ctypes.c_int64(0)
        nat.check(L.fa_sell_count(n, nat.ptr(crow), nat.ptr(self.slice_ptr), ctypes.byref(ent), nat.ptr(ws),
                                  ws.numel(), nat.stream()), "fa_sell_count")
        self.entries = int(ent.value)
        self.cols = torch.zeros(max(self.entries, 1), dtype=torch.int32, device=dev)
        if self.entries > 0:
            nat.check(L.fa_sell_fill(nat.FA_F64, nat.FA_F64, n, self.shape[1], nat.ptr(crow), nat.ptr(col), None,
                                     nat.ptr(self.slice_ptr), nat.ptr(self.cols), None, nat.stream()),
                      "fa_sell_fill")
        self.vals = {}

    def set_values(self, values, dtype=None):
        """Copy CSR values (FP64 or FP32) into the SELL layout as `dtype`."""
        dtype = values.dtype if dtype is None else dtype
        out = self.vals.get(dtype)
        if out is None:
            out = torch.zeros(max(self.entries, 1), dtype=dtype, device=values.device)
            self.vals[dtype] = out
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
        S = (key, SellMatrix(A.crow, A.col, A.shape))
        A._sell = S
    sm = S[1]
    sm.set_values(A.values)
    return sm
