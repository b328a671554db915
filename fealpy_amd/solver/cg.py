"""Conjugate gradients (fealpy/solver/cg.py:14-123).

``cg(A, b)`` with a device CSR matrix runs entirely on the GPU (``fa_pcg``:
fused SpMV + dot, update + dot, direction update; the host looks at the
convergence flag every ``check_every`` iterations).  Any other operator with
``__matmul__`` (and any preconditioner ``M``) goes through the same iteration
written with the HIP vector kernels.  A 2-D ``b`` holds one right-hand side
per column (per row with ``batch_first``); the columns iterate in lockstep as
in the reference.  Stopping rule, return values and the ``info`` dict follow
the reference.
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
        bt = b.contiguous() if batch_first else b.t().contiguous()
        x0t = None if x0 is None else (x0.contiguous() if batch_first else x0.t().contiguous())
        if M is not None and hasattr(M, 'ensure_setup') and isinstance(A, CSRTensor):
            M.ensure_setup(A)
        xt, info = _cg_batched(A, bt, x0t, M, atol, rtol, maxit)
        x = xt if batch_first else xt.t().contiguous()
        return (x, info) if returninfo else x
    if M is not None and hasattr(M, 'pcg') and isinstance(A, CSRTensor):
        x, info = M.pcg(A, b, x0, atol=atol, rtol=rtol, maxit=maxit)
    elif M is None and isinstance(A, CSRTensor):
        x, info = _cg_native(A, b, x0, atol, rtol, maxit, check_every)
    else:
        x, info = _cg_generic(A, b, x0, M, atol, rtol, maxit)
    return (x, info) if returninfo else x


def maxit_arg(maxit):
    """maxit for fa_pcg: None -> no cap (-1); 0 stops after one iteration, as
    the reference loop tests niter >= maxit after its first update."""
    return -1 if maxit is None else max(int(maxit), 1)


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
                  maxit_arg(maxit), int(check_every), ctypes.byref(niter), ctypes.byref(res),
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


def _cg_batched(A, bt, x0t, M, atol, rtol, maxit):
    """Right-hand sides bt[k] (rows of a (B, n) array) in lockstep, as
    fealpy/solver/cg.py:69-123 runs the columns of b: alpha and beta per
    column, one stopping test on sqrt(sum_k r_k . z_k) against atol and
    rtol * ||b||_F.  Every product and vector operation is a HIP kernel on one
    contiguous row."""
    info = {'residual': 0.0, 'niter': 0}
    nb = bt.shape[0]
    V = _Vec(bt)
    b_norm = sum(V.dot(bt[k], bt[k]) for k in range(nb)) ** 0.5
    if b_norm < 1e-15:
        return torch.zeros_like(bt), info
    X = torch.zeros_like(bt) if x0t is None else x0t.clone()
    R = bt.clone()
    for k in range(nb):
        V.axpby(-1.0, (A @ X[k]).contiguous(), 1.0, R[k])
    pre = (lambda v: (M @ v).contiguous()) if M is not None else (lambda v: v)
    Z = [pre(R[k]) for k in range(nb)]
    P = torch.stack(Z)
    rz = [V.dot(R[k], Z[k]) for k in range(nb)]
    n_iter = 0
    while True:
        for k in range(nb):
            Ap = (A @ P[k]).contiguous()
            alpha = rz[k] / V.dot(P[k], Ap)
            V.axpby(alpha, P[k], 1.0, X[k])
            V.axpby(-alpha, Ap, 1.0, R[k])
        Z = [pre(R[k]) for k in range(nb)]
        rz_new = [V.dot(R[k], Z[k]) for k in range(nb)]
        r_norm = max(sum(rz_new), 0.0) ** 0.5
        n_iter += 1
        info['residual'], info['niter'] = r_norm, n_iter
        if r_norm < atol or r_norm < rtol * b_norm or (maxit is not None and n_iter >= maxit):
            break
        for k in range(nb):
            V.axpby(1.0, Z[k], rz_new[k] / rz[k], P[k])
        rz = rz_new
    return X, info
