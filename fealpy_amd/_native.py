"""ctypes binding of libfealpy_amd.so (declared in include/fealpy_amd.h).

The product path has no CPU fallback: every numerical routine of the package
calls into this library, and ``lib()`` raises if the library is missing or no
GPU is visible.  torch is imported first so the library binds to the same HIP
runtime (libamdhip64.so.7) torch already loaded.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfealpy_amd.so")

FA_OK, FA_EINVAL, FA_ENOMEM, FA_EHIP, FA_ENOCONV, FA_ECAPACITY = range(6)
FA_F64, FA_F32 = 0, 1
(TIMER_PCG_SPMV, TIMER_P1_ROWS, TIMER_P1_SUM_ROWS, TIMER_P1_VECTORS, TIMER_P1_SUM_VALUES, TIMER_BC_APPLY,
 TIMER_SELL_SPMV, TIMER_VCYCLE, TIMER_P1_APPLY, TIMER_SMOOTH0, TIMER_P1_BOX) = range(1, 12)
COEF_NONE, COEF_CONST, COEF_CELL, COEF_CELLQ, COEF_SEP = 0, 1, 2, 3, 4

_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_sz = ctypes.c_size_t
_pi32 = ctypes.POINTER(ctypes.c_int32)
_pi64 = ctypes.POINTER(ctypes.c_int64)
_pf64 = ctypes.POINTER(ctypes.c_double)

_SIGNATURES = {
    "fa_last_error": (ctypes.c_char_p, []),
    "fa_version": (ctypes.c_int, []),
    "fa_adjacency_work_bytes": (_sz, [_i64, _i32, _i64]),
    "fa_adjacency_build": (ctypes.c_int, [_p, _i64, _i32, _i64, _p, _p, _pi32, _p, _sz, _p]),
    "fa_pattern_work_bytes": (_sz, [_i64]),
    "fa_pattern_count": (ctypes.c_int, [_p, _i64, _i32, _i64, _p, _p, _i32, _p, _pi64, _pi32, _p, _sz, _p]),
    "fa_pattern_fill": (ctypes.c_int, [_p, _i64, _i32, _i64, _p, _p, _i32, _p, _p, _p]),
    "fa_element_matrices": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _p, _p, _i64,
                                           _p, _p, _p, ctypes.c_int, _f64, _p, ctypes.c_int, _f64, _p, _p, _p]),
    "fa_element_vectors": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _p, _p, _i64,
                                          _p, _p, ctypes.c_int, _f64, _p, _p, _p]),
    "fa_gather_matrix": (ctypes.c_int, [ctypes.c_int, _i64, _i32, _p, _p, _p, _p, _p, _i32, _p, _i64, _p, _p]),
    "fa_gather_vector": (ctypes.c_int, [ctypes.c_int, _i64, _i32, _p, _p, _p, _p, _p]),
    "fa_element_apply": (ctypes.c_int, [ctypes.c_int, _i64, _i32, _p, _p, _i64, _p, _p, _p]),
    "fa_element_elasticity": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _p, _p, _i64, _p,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_int, _p, _p]),
    "fa_element_vectors_tensor": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, _p, _p, _i64, _p, _p, _p, _p, _p]),
    "fa_sell16_fill": (ctypes.c_int, [ctypes.c_int, _i64, _p, _p, _p, _p, _p, _p, _p, ctypes.c_size_t, _p]),
    "fa_mfma_tiles_per_group": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "fa_mfma_work_bytes": (ctypes.c_size_t, [ctypes.c_int, _i64, ctypes.c_int]),
    "fa_element_matrices_mfma": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, _p, _p, _i64, _p, _p, ctypes.c_int, ctypes.c_double, _p,
                                                ctypes.c_int, ctypes.c_double, _p, _p, _i64, _p, ctypes.c_size_t,
                                                _p]),
    "fa_assemble_p1_matrix": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i64, ctypes.c_int, _p, _p, _p, _p,
                                             ctypes.c_int, ctypes.c_double, _p, ctypes.c_int, ctypes.c_double, _p,
                                             _p, _p, _p, _p, _i32, _p, _p]),
    "fa_assemble_p1_vector": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i64, ctypes.c_int, _p, _p, _p, _p,
                                             ctypes.c_int, ctypes.c_double, _p, _p, _p, _p, _p]),
    "fa_bc_to_point": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i64, ctypes.c_int, _p, _p, _p, _p, _p]),
    "fa_dirichlet_work_bytes": (_sz, [_i64]),
    "fa_dirichlet_count": (ctypes.c_int, [_i64, _p, _p, _p, _p, _pi64, _p, _sz, _p]),
    "fa_dirichlet_fill": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i64, _p, _p, _p, _p, _p, _p, _p, _p]),
    "fa_csr_residual_work_bytes": (_sz, []),
    "fa_csr_residual": (ctypes.c_int, [ctypes.c_int, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _sz, _p]),
    "fa_dirichlet_rhs": (ctypes.c_int, [ctypes.c_int, _i64, _p, _p, _p, _p, _p, _p, _p, _p]),
    "fa_dirichlet_apply": (ctypes.c_int, [ctypes.c_int, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "fa_dirichlet_apply_keep": (ctypes.c_int, [ctypes.c_int, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "fa_spmv": (ctypes.c_int, [ctypes.c_int, _i64, _i64, _p, _p, _p, _p, _p, _p]),
    "fa_spmm": (ctypes.c_int, [ctypes.c_int, _i64, _i64, _i64, _p, _p, _p, _p, _p, _p]),
    "fa_reduce_work_bytes": (_sz, []),
    "fa_dot": (ctypes.c_int, [ctypes.c_int, _i64, _p, _p, _p, _p, _sz, _p]),
    "fa_axpby": (ctypes.c_int, [ctypes.c_int, _i64, _f64, _p, _f64, _p, _p]),
    "fa_cast": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i64, _p, _p, _p]),
    "fa_sell_work_bytes": (_sz, [_i64]),
    "fa_sell_count": (ctypes.c_int, [_i64, _p, _p, _pi64, _p, _sz, _p]),
    "fa_sell_fill": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i64, _i64, _p, _p, _p, _p, _p, _p, _p]),
    "fa_sell_spmv": (ctypes.c_int, [ctypes.c_int, _p, _p, _p, _p]),
    "fa_rap_work_bytes": (_sz, [_i64]),
    "fa_rap_count": (ctypes.c_int, [_i64, _i64, _p, _p, _p, _p, _p, _p, _i32, _p, _pi64, _pi32, _p, _sz, _p]),
    "fa_rap_fill": (ctypes.c_int, [_i64, _p, _p, _p, _p, _p, _p, _i32, _p, _p, _p]),
    "fa_rap_values": (ctypes.c_int, [ctypes.c_int, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i32, _p,
                                     _p]),
    "fa_rap_terms_work_bytes": (_sz, [_i64]),
    "fa_rap_terms_count": (ctypes.c_int, [_i64, _p, _p, _p, _p, _p, _p, _p, _p, _i32, _p, _i64, _p, ctypes.c_int, _p,
                                          _pi64, _p, _sz, _p]),
    "fa_rap_terms_fill": (ctypes.c_int, [_i64, _p, _p, _p, _p, _p, _p, _p, _p, _i32, _p, _i64, _p, ctypes.c_int, _p,
                                         _i64, _p, _p, _sz, _p]),
    "fa_rap_values_terms": (ctypes.c_int, [ctypes.c_int, _i64, _p, _p, _p, ctypes.c_int, _p, _p, _p]),
    "fa_rap_terms_fill_staged": (ctypes.c_int, [_i64, _p, _p, _p, _p, _p, _p, _p, _p, _i32, _p, _i64, _p, ctypes.c_int,
                                                _p, _i64, _p, _p, _sz, _p]),
    "fa_rap_values_staged": (ctypes.c_int, [ctypes.c_int, _i64, _p, _p, _p, _p, _p, _p, _p, ctypes.c_int, _p, _p, _p]),
    "fa_csr_to_dense": (ctypes.c_int, [_i64, _p, _p, _p, _p, _p]),
    "fa_dense_spd_inverse": (ctypes.c_int, [_i64, _p, _p]),
    "fa_mg_lmax": (ctypes.c_int, [_p, ctypes.c_int, ctypes.c_int, _pf64, _p, _sz, _p]),
    "fa_mg_vcycle": (ctypes.c_int, [_p, _p, _p, _p]),
    "fa_incidence_maps": (ctypes.c_int, [_i64, _i32, _p, _p, _p, _p, _p, _p, _p, _p, _sz, _p]),
    "fa_p1_incidence_rows": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i64, ctypes.c_int, _p, _p, _p, _p,
                                            ctypes.c_int, _f64, _p, ctypes.c_int, _f64, _p, _p, _p, _p]),
    "fa_sum_incidence_rows": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64, _p, _p, _p, _p, _i32,
                                             _p, _p]),
    "fa_p1_box_matrix": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _p, _p, ctypes.c_int,
                                        _p, _p, ctypes.c_int, _f64, _p, ctypes.c_int, _f64, _p, _p, _i64, _p, _p]),
    "fa_sep_points": (ctypes.c_int, [_p, ctypes.c_int64, ctypes.c_int, _p, _p, _p]),
    "fa_box_axes_work_bytes": (_sz, []),
    "fa_box_axes": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _p, _p, _p, _p, _sz, _p]),
    "fa_p1_box_vector_work_bytes": (_sz, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "fa_p1_box_pattern_work_bytes": (_sz, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "fa_pcg_wg_batched": (ctypes.c_int, [ctypes.c_int, _p, _p, _p]),
    "fa_p1_box_pattern": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _p, _p, _i64, _p, _sz, _p]),
    "fa_p1_box_vector": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _p, ctypes.c_int, _p,
                                        _p, ctypes.c_int, _f64, _p, _p, _p, _sz, _p]),
    "fa_p1_incidence_vectors": (ctypes.c_int, [ctypes.c_int, _i64, ctypes.c_int, _p, _p, _p, _p, ctypes.c_int, _f64,
                                               _p, _p, _p, _p]),
    "fa_sum_incidence_values": (ctypes.c_int, [ctypes.c_int, _i64, _p, _p, _p, _p]),
    "fa_refine_update": (ctypes.c_int, [ctypes.c_int, _i64, _p, _p, _p, _p]),
    "fa_tune_set": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "fa_pack_sell": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _p, _p, _i64, _i64, _p, _p, _p, _p]),
    "fa_copy_segments": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _p, _p, _p]),
    "fa_pcg_batched_work_bytes": (_sz, [_i64, ctypes.c_int]),
    "fa_pcg_batched": (ctypes.c_int, [_p, _p, ctypes.c_int, _p, _p, _p, _p, _f64, _f64, _i64, _i32, _p, _p, _p, _p,
                                      _sz, _p]),
    "fa_amg_aggregate_work_bytes": (_sz, [_i64, _i64]),
    "fa_amg_aggregate": (ctypes.c_int, [ctypes.c_int, _i64, _p, _p, _p, _f64, _p, _pi64, _pi32, _p, _sz, _p]),
    "fa_amg_smooth_prolongation": (ctypes.c_int, [ctypes.c_int, _i64, _p, _p, _p, _p, _p, _p, _p, _f64, _p, _p]),
    "fa_tune_get": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "fa_tune_key": (ctypes.c_int, [ctypes.c_int, _i64, _i64]),
    "fa_topology_work_bytes": (_sz, [_i64, _i64]),
    "fa_topology_sort": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _p, ctypes.c_int, _p, _i64, _i64, _pi64, _p, _sz,
                                        _p]),
    "fa_topology_fill": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _p, ctypes.c_int, _p, _i64, _i64, _p, _p, _p, _p,
                                        _sz, _p]),
    "fa_p1_apply_work_bytes": (_sz, []),
    "fa_p1_apply": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i64, ctypes.c_int, _p, _p, _p, _p, ctypes.c_int,
                                   _f64, _p,
                                   ctypes.c_int, _f64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _sz, _p]),
    "fa_smoother_setup": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _p, _p, _p, _p, _p, _p, _p, _sz, _p]),
    "fa_jacobi_setup": (ctypes.c_int, [ctypes.c_int, _i64, _i64, _p, _p, _p, _p, _pf64, _p, _sz, _p]),
    "fa_pcg_work_bytes": (_sz, [ctypes.c_int, _i64]),
    "fa_pcg": (ctypes.c_int, [ctypes.c_int, _p, _p, _p, _p, _f64, _f64, _i64, _i32, _pi64, _pf64, _p, _sz, _p]),
    "fa_pcg_from_zero": (ctypes.c_int, [ctypes.c_int, _p, _p, _p, _p, _f64, _f64, _i64, _i32, _pi64, _pf64, _p, _sz,
                                        _p]),
    "fa_pcg_graph_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "fa_pcg_graph_destroy": (None, [_p]),
    "fa_pcg_static": (ctypes.c_int, [ctypes.c_int, _p, _p, _p, _p, _f64, _f64, _i64, _i32, _pi64, _pf64, _p]),
    "fa_timer_arm": (ctypes.c_int, [ctypes.c_uint]),
    "fa_timer_read": (ctypes.c_int, [ctypes.c_int, _pi64, _pf64]),
    "fa_pcg_graph": (ctypes.c_int, [ctypes.c_int, _p, _p, _p, _p, _f64, _f64, _i64, _i32, _pi64, _pf64, _p, _sz, _p,
                                    _p]),
}

_lib = None


class NativeError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load the shared library and declare every C-ABI signature (no GPU needed)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise NativeError(f"{path} is missing: build it with `python -m fealpy_amd.build`")
        lib = ctypes.CDLL(path)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def lib():
    """The library, for compute calls: requires a visible GPU."""
    if not torch.cuda.is_available():
        raise NativeError("fealpy_amd needs a ROCm GPU (MI355X): torch.cuda.is_available() is False")
    return load()


def check(status, what=""):
    if status == FA_OK:
        return
    msg = _lib.fa_last_error().decode() if _lib is not None else ""
    if status == FA_EINVAL:
        raise ValueError(f"{what}: {msg}")
    if status == FA_ENOMEM:
        raise MemoryError(f"{what}: {msg}")
    raise NativeError(f"{what} failed (status {status}): {msg}")


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("fealpy_amd kernels need device (cuda) tensors")
    return ctypes.c_void_p(t.data_ptr())


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_get_device = getattr(torch._C, "_cuda_getDevice", None)


def stream():
    """The calling thread's current HIP stream of the current device (the
    stream torch's own kernels use).  Read through torch's raw accessors when
    it has them: the public torch.cuda.current_stream() resolves the device
    in Python, about 9 us a call, and a small solve makes ~25 library calls."""
    if _raw_stream is not None and _get_device is not None:
        return ctypes.c_void_p(_raw_stream(_get_device()))
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def dtype_code(dt):
    if dt == torch.float64:
        return FA_F64
    if dt == torch.float32:
        return FA_F32
    raise ValueError(f"unsupported value dtype {dt}; use float64 or float32")


def workspace(nbytes, device):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def timer_arm(*kinds):
    """Arm the in-step kernel timer for the FA_TIMER_* kinds (none: disarm)."""
    mask = 0
    for k in kinds:
        mask |= 1 << int(k)
    check(lib().fa_timer_arm(mask), "fa_timer_arm")


def timer_read(kind):
    """(launches, total ms) of one armed kind since the last arming."""
    n = ctypes.c_int64(0)
    ms = ctypes.c_double(0.0)
    check(lib().fa_timer_read(int(kind), ctypes.byref(n), ctypes.byref(ms)), "fa_timer_read")
    return int(n.value), float(ms.value)
