"""Conjugate gradients (fealpy/solver/cg.py:14-123).

``cg(A, b)`` with a device CSR matrix runs entirely on the GPU (``fa_cg``:
fused SpMV + dot, update + dot, direction update; the host looks at the
convergence flag every ``check_every`` iterations).  Any other operator with
``__matmul__`` (and any preconditioner ``M``) goes through the same iteration
written with the HIP vector kernels.  Stopping rule, return values and the
``info`` dict follow the reference.
"""
import torch

from .. import _native as nat
from ..sparse.tensors import CSRTensor

__all__ = ["cg"]


def cg(A, b, x0=None, M=None, *, batch_first=False, atol=1e-12, rtol=1e-8, maxit=10000, returninfo=False,
       check_every=16):
    if not torch.is_tensor(b):
        raise TypeError("b must be a Tensor")
    if b.ndim not in (1, 2):
        raise ValueError("b must be a 1D or 2D dense tensor")
    if x0 is not None and x0.shape != b.shape:
        raise ValueError("x0 and b must have the same shape")
    if b.ndim == 2:
        raise NotImplementedError("multiple right-hand sides: solve them one by one or use fealpy_amd.serving")
    if M is not None and hasattr(M, 'pcg') and isinstance(A, CSRTensor):
        x, info = M.pcg(A, b, x0, atol=atol, rtol=rtol, maxit=maxit)
    elif M is None and isinstance(A, CSRTensor):
        x, info = _cg_native(A, b, x0, atol, rtol, maxit, check_every)
    else:
        x, info = _cg_generic(A, b, x0, M, atol, rtol, maxit)
    return (x, info) if returninfo else x


def _cg_native(A, b, x0, atol, rtol, maxit, check_every):
    """Plain CG on the SELL copy of A (fa_pcg with no preconditioner)."""
    import ctypes
    from ..sparse.sell import sell_of
    L = nat.lib()
    dt = A.dtype
    S = sell_of(A)
    b = b.to(dt).contiguous()
    x = torch.zeros_like(b) if x0 is None else x0.to(dt).clone().contiguous()
    n = b.shape[0]
    ws = nat.workspace(L.fa_pcg_work_bytes(nat.dtype_code(dt), n), b.device)
    niter = ctypes.c_int64(0)
    res = ctypes.c_double(0.0)
    d = S.desc(dt)
    st = L.fa_pcg(nat.dtype_code(dt), ctypes.byref(d), None, nat.ptr(b), nat.ptr(x), float(atol), float(rtol),
                  int(maxit if maxit is not None else 0), int(check_every), ctypes.byref(niter), ctypes.byref(res),
                  nat.ptr(ws), ws.numel(), nat.stream())
    if st != nat.FA_ENOCONV:
        nat.check(st, "fa_pcg")
    return x, {'residual': float(res.value), 'niter': int(niter.value)}


class _Vec:
    """Thin helpers over the HIP vector kernels."""

    def __init__(self, like):
        self.L = nat.lib()
        self.code = nat.dtype_code(like.dtype)
        self.out = torch.empty(1, dtype=torch.float64, device=like.device)
        self.ws = nat.workspace(self.L.fa_reduce_work_bytes(), like.device)

    def dot(self, x, y):
        nat.check(self.L.fa_dot(self.code, x.numel(), nat.ptr(x), nat.ptr(y), nat.ptr(self.out), nat.ptr(self.ws),
                                self.ws.numel(), nat.stream()), "fa_dot")
        return float(self.out.item())

    def axpby(self, a, x, b, y):
        nat.check(self.L.fa_axpby(self.code, x.numel(), float(a), nat.ptr(x), float(b), nat.ptr(y), nat.stream()),
                  "fa_axpby")
        return y


def _cg_generic(A, b, x0, M, atol, rtol, maxit):
    info = {'residual': 0.0, 'niter': 0}
    b = b.contiguous()
    V = _Vec(b)
    b_norm = V.dot(b, b) ** 0.5
    if b_norm < 1e-15:
        return torch.zeros_like(b), info
    x = torch.zeros_like(b) if x0 is None else x0.clone().contiguous()
    r = b.clone()
    V.axpby(-1.0, (A @ x).contiguous(), 1.0, r)
    z = (M @ r).contiguous() if M is not None else r
    p = z.clone()
    rz = V.dot(r, z)
    n_iter = 0
    while True:
        Ap = (A @ p).contiguous()
        alpha = rz / V.dot(p, Ap)
        V.axpby(alpha, p, 1.0, x)
        V.axpby(-alpha, Ap, 1.0, r)
        z = (M @ r).contiguous() if M is not None else r
        rz_new = V.dot(r, z)
        r_norm = max(rz_new, 0.0) ** 0.5
        n_iter += 1
        info['residual'], info['niter'] = r_norm, n_iter
        if r_norm < atol or r_norm < rtol * b_norm or (maxit is not None and n_iter >= maxit):
            break
        V.axpby(1.0, z, rz_new / rz, p)
        rz = rz_new
    return x, info
