"""Mixed-precision solve: FP32 inner MG-PCG, FP64 residuals (iterative refinement).

The reference solves its systems in FP64 throughout (fealpy/solver/cg.py:14-123
on the matrix of BilinearForm.assembly, fem/bilinear_form.py:83-105, after
DirichletBC.apply, fem/dirichlet_bc.py:101-121).  ``cg_refine`` reaches the
same FP64 solution with the bulk of the work in FP32:

- the inner operator is an FP32 Dirichlet-applied matrix (assembled in FP32,
  or filtered from the FP64 matrix with its values rounded in the same pass)
  and the V-cycle is set up from it;
- the residual b - A x of the FP64 system is formed in FP64, either from the
  assembled FP64 matrix before the Dirichlet condition (one SpMV pass,
  fa_csr_residual) or matrix-free (``BilinearForm.residual``: element rows
  recomputed from the FP64 geometry and coefficients, no FP64 matrix at all);
- each correction A32 d = r is an FP32 MG-PCG solve to ``inner_rtol``, added
  to the FP64 iterate off the Dirichlet rows.

The Dirichlet condition enters as in the reference: boundary DOFs carry the
interpolated boundary values, and on the other rows the residual of the
Dirichlet-applied system (b - A g on the free rows, the identity on the
boundary rows) equals b - A x of the unmodified system whenever x equals g
on the boundary, which the correction step keeps exactly.

The stopping rule is the reference cg's, on the FP64 residual: ||r|| < atol or
||r|| < rtol ||b'||, with b' the Dirichlet-applied right-hand side.
"""
import torch

from .. import _native as nat
from ..sparse.tensors import CSRTensor

__all__ = ["cg_refine"]


def _residual(op, b, x, bd, dtype):
    """(b - A x off the Dirichlet rows, 0 on them, as `dtype`; its squared
    norm as a 1-element FP64 device tensor) for an FP64 CSR matrix or a
    BilinearForm (matrix-free)."""
    if not isinstance(op, CSRTensor):
        return op.residual(b, x, bd, dtype=dtype)
    L = nat.lib()
    n = b.shape[0]
    y = torch.empty(n, dtype=dtype, device=b.device)
    norm2 = torch.empty(1, dtype=torch.float64, device=b.device)
    ws = nat.workspace(L.fa_csr_residual_work_bytes(), b.device)
    nat.check(L.fa_csr_residual(nat.dtype_code(dtype), n, nat.ptr(op.crow), nat.ptr(op.col), nat.ptr(op.values),
                                nat.ptr(x), nat.ptr(b), nat.ptr(bd), nat.ptr(y), nat.ptr(norm2), nat.ptr(ws),
                                ws.numel(), nat.stream()), "fa_csr_residual")
    return y, norm2


def cg_refine(op, b, A, M, bc, *, x0=None, atol=0.0, rtol=1e-8, inner_rtol=1e-3, maxit=30, inner_maxit=1000,
              inner_check=8, returninfo=False):
    """Solve the FP64 Dirichlet problem of the operator ``op`` and the load
    vector ``b`` (FP64, before the boundary condition) by iterative refinement.

    op: the FP64 operator before the Dirichlet condition: its assembled FP64
        CSRTensor, or the BilinearForm itself (applied matrix-free);
    A:  the Dirichlet-applied matrix of the inner solves (any precision; FP32
        for the mixed form), e.g. ``bc.apply_matrix(A64, dtype=torch.float32)``;
    M:  preconditioner of the inner solves (a GeometricMG set up for A);
    bc: the DirichletBC (boundary flags and values).

    inner_check: the inner solves look at their convergence flag at most every
    this many iterations (one host synchronisation each), the first time after
    the iteration count of the last inner solve to the same ``inner_rtol``
    with ``M``; the iterations past convergence before a look return at once
    on the device.

    Returns x (FP64), and with ``returninfo`` a dict with the final FP64
    residual norm, the inner PCG iterations in total (``niter``) and the
    number of corrections (``refinements``)."""
    L = nat.lib()
    n = bc.gdof
    if isinstance(op, CSRTensor):
        if op.dtype != torch.float64 or op.shape != (n, n):
            raise ValueError("op must be the FP64 matrix of the space (before the Dirichlet condition)")
    elif op._spaces[0].number_of_global_dofs() != n:
        raise ValueError("the form's space does not match the boundary condition's")
    if b.shape != (n,):
        raise ValueError("b must be a vector over the space's DOFs")
    if A.shape != (n, n):
        raise ValueError("A does not match the space")
    gd = bc.gd
    if gd is None:
        raise RuntimeError("The boundary condition is None.")
    dev = b.device
    bd = bc._bd_u8
    b = b.to(torch.float64).contiguous()
    g = bc._boundary_values(b, None, gd, torch.float64)          # g on the boundary, 0 elsewhere
    ws = nat.workspace(L.fa_reduce_work_bytes(), dev)
    gg = torch.empty(1, dtype=torch.float64, device=dev)
    nat.check(L.fa_dot(nat.FA_F64, n, nat.ptr(g), nat.ptr(g), nat.ptr(gg), nat.ptr(ws), ws.numel(), nat.stream()),
              "fa_dot")
    inner_dt = A.dtype
    # b' = (b - A g) off the boundary, g on it: ||b'||^2 = ||b - A g||^2 (free rows) + ||g||^2
    r, nr2 = _residual(op, b, g, bd, inner_dt)
    bnorm2 = nr2 + gg
    x = g
    if x0 is not None:                      # x = g on the boundary rows, x0 elsewhere
        x = g.clone()
        nat.check(L.fa_refine_update(nat.FA_F64, n, nat.ptr(x), nat.ptr(x0.to(torch.float64).contiguous()),
                                     nat.ptr(bd), nat.stream()), "fa_refine_update")
        r, nr2 = _residual(op, b, x, bd, inner_dt)
    hb = torch.cat([nr2, bnorm2]).tolist()
    rnorm, bnorm = hb[0] ** 0.5, hb[1] ** 0.5
    info = {'residual': rnorm, 'niter': 0, 'refinements': 0}
    if bnorm < 1e-15:
        info['residual'] = 0.0
        x.zero_()
        return (x, info) if returninfo else x
    tol = max(atol, rtol * bnorm)
    # The inner solves first look at their flag after as many iterations as
    # the last inner solve to the same tolerance with this preconditioner took
    # (at most inner_check): each look past convergence costs the iterations
    # launched before it, which return at once but still cost their launches.
    # The iterates and counts do not depend on when the host looks.
    hints = M.__dict__.setdefault('_inner_niter_hint', {})
    while rnorm >= tol and info['refinements'] < maxit:
        check = max(1, min(inner_check, hints.get(inner_rtol, inner_check)))
        d, inner = M.pcg(A, r, atol=0.0, rtol=inner_rtol, maxit=inner_maxit, check_every=check)
        if inner['niter'] > 0:
            hints[inner_rtol] = inner['niter']
        nat.check(L.fa_refine_update(nat.dtype_code(d.dtype), n, nat.ptr(x), nat.ptr(d), nat.ptr(bd), nat.stream()),
                  "fa_refine_update")
        info['niter'] += inner['niter']
        info['refinements'] += 1
        r, nr2 = _residual(op, b, x, bd, inner_dt)
        rnorm = float(nr2.item()) ** 0.5
        info['residual'] = rnorm
    return (x, info) if returninfo else x
