"""ctypes binding of fealpy's sparse-product hook to libfealpy_amd.so.

This is the file a fealpy maintainer would add as
``fealpy/backend/mi355x.py``.  It depends on torch and the shared library
only (not on the rest of fealpy_amd), so it can be copied as it is.

fealpy routes every ``CSRTensor.matmul`` through the backend's
``csr_spmm(crow, col, values, shape, other)`` (fealpy/sparse/csr_tensor.py:
411-452).  On the torch backend that is ``PyTorchBackend.csr_spmm``
(fealpy/backend/pytorch_backend.py:527-533), a staticmethod defined on the
class itself; ``ModuleProxy.attach_methods`` (fealpy/backend/base.py:250-263)
skips names a backend already defines, so the hook is installed by
assigning the staticmethod directly::

    from fealpy.backend.pytorch_backend import PyTorchBackend
    from fealpy.backend import mi355x
    mi355x.register(PyTorchBackend)

Semantics follow the numpy backend's csr_spmm (numpy_backend.py:181-199):
``other`` 1-D -> y (M,), 2-D (N, K) -> Y (M, K) (SciPy's csr_matvecs
branch); other ranks raise ValueError; batched values (values.ndim != 1)
raise NotImplementedError.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("FEALPY_AMD_LIB", os.path.join(_HERE, "libfealpy_amd.so"))
_F64, _F32 = 0, 1
_lib = None


def _load():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(LIB_PATH)       # torch is imported first: one libamdhip64 for both
        p, i64 = ctypes.c_void_p, ctypes.c_int64
        lib.fa_spmv.restype = ctypes.c_int
        lib.fa_spmv.argtypes = [ctypes.c_int, i64, i64, p, p, p, p, p, p]
        lib.fa_spmm.restype = ctypes.c_int
        lib.fa_spmm.argtypes = [ctypes.c_int, i64, i64, i64, p, p, p, p, p, p]
        lib.fa_last_error.restype = ctypes.c_char_p
        _lib = lib
    return _lib


def csr_spmm(crow, col, values, shape, other):
    """y = A @ other on the GPU (fa_spmv / fa_spmm); drop-in for
    PyTorchBackend.csr_spmm with the numpy backend's 1-D / 2-D semantics."""
    if values.ndim != 1:
        raise NotImplementedError("Batch sparse matrix multiplication has not been supported yet.")
    if other.ndim not in (1, 2):
        raise ValueError("`other` must be a 1-D or 2-D array.")
    if values.dtype not in (torch.float64, torch.float32):
        raise TypeError(f"values of dtype {values.dtype}: float64 or float32 expected")
    if not (values.is_cuda and other.is_cuda):
        raise ValueError("csr_spmm on the MI355X needs device tensors")
    M, N = int(shape[0]), int(shape[1])
    if other.shape[0] != N:
        raise ValueError(f"`other` has {other.shape[0]} rows for a matrix with {N} columns")
    lib = _load()
    crow32 = crow.to(torch.int32).contiguous()
    col32 = col.to(torch.int32).contiguous()
    val = values.contiguous()
    x = other.to(val.dtype).contiguous()
    code = _F64 if val.dtype == torch.float64 else _F32
    stream = ctypes.c_void_p(torch.cuda.current_stream(val.device).cuda_stream)
    if x.ndim == 1:
        y = torch.empty(M, dtype=val.dtype, device=val.device)
        st = lib.fa_spmv(code, M, col32.numel(), crow32.data_ptr(), col32.data_ptr(), val.data_ptr(), x.data_ptr(),
                         y.data_ptr(), stream)
    else:
        y = torch.empty((M, x.shape[1]), dtype=val.dtype, device=val.device)
        st = lib.fa_spmm(code, M, col32.numel(), x.shape[1], crow32.data_ptr(), col32.data_ptr(), val.data_ptr(),
                         x.data_ptr(), y.data_ptr(), stream) if x.shape[1] else 0
    if st != 0:
        raise RuntimeError(f"csr_spmm: {lib.fa_last_error().decode()}")
    return y


def register(backend_cls):
    """Install csr_spmm on a fealpy backend class (PyTorchBackend): a direct
    staticmethod assignment, since attach_methods keeps methods a backend
    defines itself."""
    backend_cls.csr_spmm = staticmethod(csr_spmm)
    return backend_cls
