"""Geometric multigrid preconditioner for P1 problems on nested ``from_box`` meshes.

Role of the reference's ``GAMGSolver`` (fealpy/solver/gamg_solver.py:66-301):
a hierarchy of operators A_l, transfers P_l / R_l = P_l^T, V-cycles used as
the preconditioner of CG.  Designed for the GPU:

* levels are the ``from_box`` meshes with every cell count halved; the Kuhn /
  diagonal-split triangulations are nested under halving, so P1 prolongation
  is closed form: a fine node with grid index g takes 1/2 of the coarse
  nodes floor(g/2) and ceil(g/2) (one node with weight 1 when g is even);
* coarse operators are Galerkin products R A P formed on the GPU
  (``fa_rap_*``: symbolic once, numeric per setup, fixed summation order);
* Dirichlet rows stay decoupled: P has no rows for fine Dirichlet nodes and
  no columns for coarse Dirichlet nodes; coarse Dirichlet rows become unit rows;
* smoothing is a Chebyshev polynomial in D^-1 A (degree ``degree``, interval
  [lower, upper] * the Gershgorin bound of D^-1 A -- an upper bound of the
  spectrum, so no mode is amplified), the same polynomial
  before and after the coarse correction, so the V-cycle is symmetric and
  CG-compatible; the coarsest level is solved with its dense inverse;
* the cycle runs in ``dtype`` (FP32 for the mixed-precision solver: FP64 CG
  with an FP32 V-cycle).
"""
import ctypes

import torch

from .. import _native as nat
from ..sparse.sell import FaSell, SellMatrix
from ..sparse.tensors import CSRTensor


class FaMgLevel(ctypes.Structure):
    _fields_ = [("A", FaSell), ("dinv", ctypes.c_void_p), ("R", FaSell), ("P", FaSell), ("lmax", ctypes.c_double),
                ("b", ctypes.c_void_p), ("x", ctypes.c_void_p), ("x2", ctypes.c_void_p), ("d", ctypes.c_void_p),
                ("r", ctypes.c_void_p)]


class FaMg(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int), ("nlevels", ctypes.c_int), ("levels", ctypes.POINTER(FaMgLevel)),
                ("degree", ctypes.c_int), ("lower", ctypes.c_double), ("upper", ctypes.c_double),
                ("coarse_chol", ctypes.c_void_p)]


def _grid_shape(mesh):
    meta = mesh.meshdata.get('from_box')
    if meta is None:
        raise ValueError("GeometricMG needs a mesh made by from_box (nested halving is closed form only there)")
    return list(meta[2:])


def _node_ids(dims, g):
    """meshgrid('ij') node id of grid indices g (..., d) for cell counts dims."""
    idx = torch.zeros(g.shape[:-1], dtype=torch.int64, device=g.device)
    for k, n in enumerate(dims):
        idx = idx * (n + 1) + g[..., k]
    return idx


def _grid_index(dims, ids):
    out = []
    for n in reversed(dims):
        out.append(ids % (n + 1))
        ids = ids // (n + 1)
    return torch.stack(out[::-1], dim=-1)


def _prolongation(dims_f, bd_f, bd_c, device):
    """CSR (int32/FP64) of P1 prolongation coarse -> fine for nested grids."""
    dims_c = [n // 2 for n in dims_f]
    nf = 1
    for n in dims_f:
        nf *= n + 1
    fine = torch.arange(nf, device=device)
    g = _grid_index(dims_f, fine)
    lo, hi = g // 2, (g + 1) // 2
    c_lo, c_hi = _node_ids(dims_c, lo), _node_ids(dims_c, hi)
    same = c_lo == c_hi
    rows = torch.cat([fine, fine[~same]])
    cols = torch.cat([c_lo, c_hi[~same]])
    vals = torch.cat([torch.where(same, 1.0, 0.5).to(torch.float64), torch.full((int((~same).sum()),), 0.5,
                                                                             dtype=torch.float64, device=device)])
    keep = ~bd_f[rows] & ~bd_c[cols]
    rows, cols, vals = rows[keep], cols[keep], vals[keep]
    order = torch.argsort(rows * (cols.max() + 1 if cols.numel() else 1) + cols)
    rows, cols, vals = rows[order], cols[order], vals[order]
    nc = 1
    for n in dims_c:
        nc *= n + 1
    return _csr(rows, cols, vals, nf, nc)


def _csr(rows, cols, vals, nrows, ncols):
    crow = torch.zeros(nrows + 1, dtype=torch.int64, device=rows.device)
    crow[1:] = torch.cumsum(torch.bincount(rows, minlength=nrows), 0)
    return CSRTensor(crow.to(torch.int32), cols.to(torch.int32), vals, (nrows, ncols))


def _transpose(M):
    rows = M.row_index()
    cols = M.col.long()
    order = torch.argsort(cols * M.shape[0] + rows)
    return _csr(cols[order], rows[order], M.values[order], M.shape[1], M.shape[0])


def _max_row(M):
    return int((M.crow[1:] - M.crow[:-1]).max().item()) if M.shape[0] else 0


class _Level:
    pass


class GeometricMG:
    def __init__(self, space, is_bd, *, dtype=torch.float32, degree=2, lower=0.1, upper=1.0, coarse_max=64,
                 power_iters=12):
        if space.p != 1:
            raise NotImplementedError("GeometricMG covers P1 spaces")
        mesh = space.mesh
        dims = _grid_shape(mesh)
        dev = space.device
        self.dtype, self.degree, self.lower, self.upper = dtype, degree, lower, upper
        self.power_iters = power_iters
        self.device = dev
        self.levels = []
        bd = is_bd.to(torch.bool)
        lev = _Level()
        lev.dims, lev.n, lev.bd = dims, space.number_of_global_dofs(), bd
        self.levels.append(lev)
        while all(n % 2 == 0 and n >= 2 for n in dims) and lev.n > coarse_max:
            dims_c = [n // 2 for n in dims]
            nc = 1
            for n in dims_c:
                nc *= n + 1
            g = _grid_index(dims_c, torch.arange(nc, device=dev))
            bd_c = bd[_node_ids(dims, 2 * g)]
            lev.P = _prolongation(dims, bd, bd_c, dev)
            lev.R = _transpose(lev.P)
            lev.Ps = SellMatrix(lev.P.crow, lev.P.col, lev.P.shape)
            lev.Rs = SellMatrix(lev.R.crow, lev.R.col, lev.R.shape)
            lev.Ps.set_values(lev.P.values, dtype)
            lev.Rs.set_values(lev.R.values, dtype)
            nxt = _Level()
            nxt.dims, nxt.n, nxt.bd = dims_c, nc, bd_c
            self.levels.append(nxt)
            lev, dims, bd = nxt, dims_c, bd_c
        if self.levels[-1].n > 2048:
            raise ValueError(f"coarsest level has {self.levels[-1].n} unknowns; the dense coarse solver holds 2048")
        for lv in self.levels:
            z = lambda: torch.zeros(lv.n, dtype=dtype, device=dev)
            lv.b, lv.x, lv.x2, lv.d, lv.r, lv.dinv = z(), z(), z(), z(), z(), z()
        self._pattern_key = None
        self._desc = None
        self._setup_A = None

    # -- per-matrix setup -----------------------------------------------------
    def _symbolic(self, A):
        from ..sparse.sell import sell_of
        L = nat.lib()
        self.levels[0].A = A
        self.levels[0].As = sell_of(A)
        for l in range(len(self.levels) - 1):
            f, c = self.levels[l], self.levels[l + 1]
            cand = _max_row(f.R) * _max_row(f.A) * max(_max_row(f.P), 1)
            cptr = torch.empty(c.n + 1, dtype=torch.int32, device=self.device)
            ws = nat.workspace(L.fa_rap_work_bytes(c.n), self.device)
            nnz, mrow = ctypes.c_int64(0), ctypes.c_int32(0)
            nat.check(L.fa_rap_count(c.n, c.n, nat.ptr(f.R.crow), nat.ptr(f.R.col), nat.ptr(f.A.crow),
                                     nat.ptr(f.A.col), nat.ptr(f.P.crow), nat.ptr(f.P.col), max(cand, 1),
                                     nat.ptr(cptr), ctypes.byref(nnz), ctypes.byref(mrow), nat.ptr(ws), ws.numel(),
                                     nat.stream()), "fa_rap_count")
            cidx = torch.empty(max(int(nnz.value), 1), dtype=torch.int32, device=self.device)
            nat.check(L.fa_rap_fill(c.n, nat.ptr(f.R.crow), nat.ptr(f.R.col), nat.ptr(f.A.crow), nat.ptr(f.A.col),
                                    nat.ptr(f.P.crow), nat.ptr(f.P.col), max(cand, 1), nat.ptr(cptr), nat.ptr(cidx),
                                    nat.stream()), "fa_rap_fill")
            c.A = CSRTensor(cptr, cidx[:int(nnz.value)], torch.zeros(int(nnz.value), dtype=torch.float64,
                                                                     device=self.device), (c.n, c.n))
            c.max_row = int(mrow.value)
            c.As = SellMatrix(c.A.crow, c.A.col, c.A.shape)
        last = self.levels[-1]
        last.dense = torch.zeros((last.n, last.n), dtype=torch.float64, device=self.device)

    def setup(self, A):
        """Numeric setup for the (Dirichlet-applied, FP64) fine matrix A."""
        if A.dtype != torch.float64:
            raise TypeError("GeometricMG.setup expects the FP64 matrix")
        key = (A.crow.data_ptr(), A.col.data_ptr(), A.nnz)
        if key != self._pattern_key:
            self._symbolic(A)
            self._pattern_key = key
        L = nat.lib()
        self.levels[0].A = A
        for l in range(len(self.levels) - 1):
            f, c = self.levels[l], self.levels[l + 1]
            nat.check(L.fa_rap_values(c.n, nat.ptr(f.R.crow), nat.ptr(f.R.col), nat.ptr(f.R.values),
                                      nat.ptr(f.A.crow), nat.ptr(f.A.col), nat.ptr(f.A.values), nat.ptr(f.P.crow),
                                      nat.ptr(f.P.col), nat.ptr(f.P.values), nat.ptr(c.A.crow), nat.ptr(c.A.col),
                                      c.max_row, nat.ptr(c.A.values), nat.stream()), "fa_rap_values")
        code = nat.dtype_code(self.dtype)
        ws = nat.workspace(L.fa_reduce_work_bytes(), self.device)
        for lv in self.levels:
            lv.As.set_values(lv.A.values, self.dtype)
            lam = ctypes.c_double(0.0)
            nat.check(L.fa_jacobi_setup(code, lv.n, lv.A.nnz, nat.ptr(lv.A.crow), nat.ptr(lv.A.col),
                                        nat.ptr(lv.A.values), nat.ptr(lv.dinv), ctypes.byref(lam), nat.ptr(ws),
                                        ws.numel(), nat.stream()), "fa_jacobi_setup")
            lv.lmax = float(lam.value)
        last = self.levels[-1]
        nat.check(L.fa_csr_to_dense(last.n, nat.ptr(last.A.crow), nat.ptr(last.A.col), nat.ptr(last.A.values),
                                    nat.ptr(last.dense), nat.stream()), "fa_csr_to_dense")
        nat.check(L.fa_dense_spd_inverse(last.n, nat.ptr(last.dense), nat.stream()), "fa_dense_spd_inverse")
        self._build_desc()
        self._setup_A = A
        return self

    def estimate_lmax(self, level, iters=None):
        """Power-iteration estimate of lambda_max(D^-1 A) on a level (diagnostics)."""
        L = nat.lib()
        ws = nat.workspace(L.fa_reduce_work_bytes(), self.device)
        lam = ctypes.c_double(0.0)
        nat.check(L.fa_mg_lmax(ctypes.byref(self._desc), level, iters or self.power_iters, ctypes.byref(lam),
                               nat.ptr(ws), ws.numel(), nat.stream()), "fa_mg_lmax")
        return float(lam.value)

    def _build_desc(self):
        n = len(self.levels)
        arr = (FaMgLevel * n)()
        empty = FaSell(0, 0, None, None, None)
        for i, lv in enumerate(self.levels):
            e = arr[i]
            e.A = lv.As.desc(self.dtype)
            e.dinv = lv.dinv.data_ptr()
            e.R = lv.Rs.desc(self.dtype) if i < n - 1 else empty
            e.P = lv.Ps.desc(self.dtype) if i < n - 1 else empty
            e.lmax = getattr(lv, 'lmax', 2.0)
            e.b, e.x, e.x2, e.d, e.r = (lv.b.data_ptr(), lv.x.data_ptr(), lv.x2.data_ptr(), lv.d.data_ptr(),
                                        lv.r.data_ptr())
        self._lv_array = arr
        self._desc = FaMg(nat.dtype_code(self.dtype), n, arr, self.degree, self.lower, self.upper,
                          self.levels[-1].dense.data_ptr())

    # -- use ----------------------------------------------------------------------
    def __matmul__(self, r):
        """z = V(r): one V-cycle (r in the cycle dtype)."""
        z = torch.empty_like(r)
        nat.check(nat.lib().fa_mg_vcycle(ctypes.byref(self._desc), nat.ptr(r.contiguous()), nat.ptr(z),
                                         nat.stream()), "fa_mg_vcycle")
        return z

    def pcg(self, A, b, x0=None, *, atol=1e-12, rtol=1e-8, maxit=10000, check_every=4):
        """Preconditioned CG on the device: A's dtype outside, V-cycle in self.dtype."""
        from ..sparse.sell import sell_of
        if self._setup_A is not A:
            self.setup(A if A.dtype == torch.float64 else A.astype(torch.float64))
            self._setup_A = A
        L = nat.lib()
        S = sell_of(A)
        dt = A.dtype
        b = b.to(dt).contiguous()
        x = torch.zeros_like(b) if x0 is None else x0.to(dt).clone().contiguous()
        ws = nat.workspace(L.fa_pcg_work_bytes(nat.dtype_code(dt), b.shape[0]), b.device)
        niter, res = ctypes.c_int64(0), ctypes.c_double(0.0)
        d = S.desc(dt)
        st = L.fa_pcg(nat.dtype_code(dt), ctypes.byref(d), ctypes.byref(self._desc), nat.ptr(b), nat.ptr(x),
                      float(atol), float(rtol), int(maxit or 0), int(check_every), ctypes.byref(niter),
                      ctypes.byref(res), nat.ptr(ws), ws.numel(), nat.stream())
        if st != nat.FA_ENOCONV:
            nat.check(st, "fa_pcg")
        return x, {'residual': float(res.value), 'niter': int(niter.value)}
