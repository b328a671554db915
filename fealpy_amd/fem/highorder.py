"""High-order (P2-P4) tetrahedral element matrices on the matrix cores.

``element_matrices_mfma`` evaluates diffusion + mass element matrices with
fa_element_matrices_mfma: a GEMM over quadrature points per metric component
on v_mfma_f32_32x32x16_bf16 (BF16 inputs, FP32 accumulation) or
v_mfma_f32_32x32x2f32 (FP32), the per-cell metric applied to the FP32
accumulators.  Output is FP32 (NC, ldof, ldof), cells padded to a multiple
of 32 values apart (a strided view).  The reference tensors
K_s[q][ij] are built here in FP64 and stored in the fragment order of the
kernel:  [group][s][tile][k-step][lane][t].
"""
import math

import torch

from .. import _native as nat

PRECISIONS = {'bf16': 0, 'fp32': 1}
_KS = {0: (1, 2, 3, 4, 6), 1: (2, 5, 10, 18, 28)}   # k-step counts instantiated in fa_element_matrices_mfma
_PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def reference_tensors(space, q):
    """(7, NQ, ldof*ldof) FP64: the 6 metric blocks of grad phi_i . grad phi_j
    in reduced gradients R' = R[..., 1:] - R[..., :1], then phi_i phi_j."""
    _, qw, phi, R = space.tables(q)
    Rp = (R[:, :, 1:] - R[:, :, :1]).to(torch.float64)              # (NQ, ldof, 3)
    nq, ldof = phi.shape
    K = torch.empty((7, nq, ldof, ldof), dtype=torch.float64, device=phi.device)
    for s, (m, n) in enumerate(_PAIRS):
        t = Rp[:, :, m, None] * Rp[:, None, :, n]
        K[s] = t if m == n else t + t.transpose(1, 2)
    K[6] = phi[:, :, None] * phi[:, None, :]
    return K.reshape(7, nq, ldof * ldof)


def mfma_tables(space, q, precision):
    """(btab, ks, tiles per group, groups) cached per (q, precision)."""
    cache = space.__dict__.setdefault('_mfma_tables', {})
    key = (int(q), precision)
    if key in cache:
        return cache[key]
    prec = PRECISIONS[precision]
    K = reference_tensors(space, q)
    _, nq, nn = K.shape
    per_lane = 8 if prec == 0 else 1
    kstep = 2 * per_lane
    need = math.ceil(nq / kstep)
    avail = _KS[prec]
    if need > avail[-1]:
        raise ValueError(f"quadrature of {nq} points exceeds the MFMA kernels ({avail[-1] * kstep} at most)")
    ks = min(k for k in avail if k >= need)          # k-steps compiled into the kernel (zero padded)
    tpg = nat.lib().fa_mfma_tiles_per_group(prec, ks)
    if tpg < 1:
        raise ValueError(f"quadrature of {nq} points does not fit the MFMA kernel's LDS slice")
    ntiles = math.ceil(nn / 32)
    groups = math.ceil(ntiles / tpg)
    btab = fragment_order(K, ks, per_lane, tpg, groups)
    btab = btab.to(torch.bfloat16 if prec == 0 else torch.float32).contiguous()
    cache[key] = (btab, ks, tpg, groups)
    return cache[key]


def fragment_order(K, ks, per_lane, tpg, groups):
    """Flat [group][s][tile][k-step][lane][t] copy of K (S, NQ, NN), zero
    padded: lane l of k-step j of tile T holds K[s][2*per_lane*j + per_lane*(l//32) + t][32*T + l%32]."""
    S, nq, nn = K.shape
    kp = ks * 2 * per_lane
    Kp = torch.zeros((S, kp, groups * tpg * 32), dtype=K.dtype, device=K.device)
    Kp[:, :nq, :nn] = K
    Kp = Kp.reshape(S, ks, 2, per_lane, groups, tpg, 32)
    return Kp.permute(4, 0, 5, 1, 2, 6, 3).reshape(-1)


def element_matrices_mfma(space, diff=None, mass=None, q=None, precision='bf16'):
    """FP32 (NC, ldof, ldof) element matrices of diffusion (coefficient spec
    ``diff``) + mass (``mass``), each a (kind, scale, tensor) triple from
    integrators.coef_spec or None, on a tetrahedral Lagrange space."""
    if space.TD != 3:
        raise NotImplementedError("the MFMA element path covers tetrahedra")
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
    L = nat.lib()
    mesh = space.mesh
    q = space.p + 3 if q is None else q
    btab, ks, tpg, groups = mfma_tables(space, q, precision)
    _, qw, _, _ = space.tables(q)
    ldof = space.number_of_local_dofs()
    cell = mesh.cell.to(torch.int32).contiguous()
    nc = cell.shape[0]
    (dk, ds, dc) = diff if diff is not None else (nat.COEF_NONE, 0.0, None)
    (mk, ms, mc) = mass if mass is not None else (nat.COEF_NONE, 0.0, None)
    prec = PRECISIONS[precision]
    # cells 32 floats apart (a whole number of 128-byte lines): every store of
    # a 32-column tile then covers full cache lines
    ld = ((ldof * ldof + 31) // 32) * 32
    buf = torch.empty((nc, ld), dtype=torch.float32, device=space.device)
    ws = nat.workspace(L.fa_mfma_work_bytes(prec, nc, ks), space.device)
    nat.check(L.fa_element_matrices_mfma(prec, ldof, qw.shape[0], ks, tpg, groups, nat.ptr(mesh.node),
                                         nat.ptr(cell), nc, nat.ptr(qw), nat.ptr(btab), dk, ds, nat.ptr(dc),
                                         mk, ms, nat.ptr(mc), nat.ptr(buf), ld, nat.ptr(ws), ws.numel(),
                                         nat.stream()), "fa_element_matrices_mfma")
    return torch.as_strided(buf, (nc, ldof, ldof), (ld, ldof, 1))


def mfma_flops(nc, nq, ldof):
    """Algorithmic FLOPs of the contraction: 2 NC NQ (6 + 1) ldof^2."""
    return 2 * nc * nq * 7 * ldof * ldof
