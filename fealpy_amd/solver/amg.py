"""Smoothed-aggregation algebraic multigrid (the algebraic branch of the
reference's GAMGSolver, fealpy/solver/gamg_solver.py:104-121, and
fealpy/solver/amg.py): a hierarchy built from the matrix alone, for meshes
that are not ``from_box`` or when no geometric transfer is wanted.

Setup per matrix pattern, on the device:

* aggregates from a distance-2 maximal independent set of the strength graph
  (``fa_amg_aggregate``; strong when |a_ij| >= theta sqrt(|a_ii a_jj|)),
  rows without a strong off-diagonal entry (Dirichlet rows) in no aggregate;
* tentative prolongation P0 (one unit entry per aggregated row) smoothed by
  one damped Jacobi step, P = (I - omega/rho D^-1 A) P0 with rho the
  Gershgorin bound of D^-1 A (``fa_amg_smooth_prolongation``; A P0 by the
  Galerkin kernels with R = I), on levels whose rows hold at most
  ``smooth_max_row`` entries on average; denser levels keep P0 (plain
  aggregation), which holds the operator complexity down where smoothed
  transfers would fill the coarse rows;
* R = P^T, coarse operators R A P by the same Galerkin kernels as the
  geometric hierarchy, until a level has at most ``coarse_max`` rows.

A new matrix on the same pattern keeps the transfers and redoes only the
Galerkin products and the smoother data (the reference's GAMGSolver also
keeps its transfers across solves of one pattern).  The cycle, the smoother
(Chebyshev in D^-1 A), the dense coarsest solve, mixed precision and the
PCG driver are those of GeometricMG.
"""
import ctypes

import torch

from .. import _native as nat
from ..sparse.sell import sell_of
from .mg import GeometricMG, _Level, _csr, _identity, _max_row


class AlgebraicMG(GeometricMG):
    def __init__(self, *, dtype=torch.float32, degree=2, lower=0.1, upper=1.0, coarse_max=64, theta=0.08,
                 omega=4.0 / 3.0, smooth_max_row=30, max_levels=20, coarse_degree=12, compress=True, graphs=False,
                 power_iters=12):
        self.dtype, self.degree, self.lower, self.upper = dtype, degree, lower, upper
        self.smooth_max_row = smooth_max_row
        self.compress = bool(compress) and dtype == torch.float32
        self.coarse_max, self.theta, self.omega, self.max_levels = coarse_max, theta, omega, max_levels
        self.coarse_degree, self.power_iters = coarse_degree, power_iters
        self.graphs = bool(graphs)
        self._graph = None
        self.device = None
        self.levels = []
        self.aggregation_rounds = []
        self._pattern_key = None
        self._desc = None
        self._setup_A = None

    def setup(self, A):
        """Hierarchy for A (FP64 or FP32, Dirichlet-applied): built on the
        first matrix of a pattern, numeric products and smoother data on every
        call."""
        key = (A.crow.data_ptr(), A.col.data_ptr(), A.nnz)
        if key != self._pattern_key:
            self._build(A)
            self._pattern_key = key
        return GeometricMG.setup(self, A)

    # -- hierarchy ----------------------------------------------------------------
    def _build(self, A):
        self.device = A.device
        lev = _Level()
        lev.n = A.shape[0]
        lev.A, lev.As, lev.max_row, lev.compressed = A, sell_of(A), _max_row(A), None
        self.levels = [lev]
        self.aggregation_rounds = []
        while self.levels[-1].n > self.coarse_max and len(self.levels) < self.max_levels:
            f = self.levels[-1]
            P = self._prolongation(f.A)
            if P is None:
                break
            self._link(f, P)
            c = _Level()
            c.n, c.compressed = P.shape[1], None
            self.levels.append(c)
            l = len(self.levels) - 2
            self._symbolic_pair(l)
            self._numeric_pair(l, f.A.dtype)
        self._finish_levels()
        last = self.levels[-1]
        last.dense = (torch.zeros((last.n, last.n), dtype=torch.float64, device=self.device) if self.coarse_dense
                      else None)

    def _prolongation(self, A):
        """Smoothed aggregation prolongation of one level, or None when the
        level does not coarsen (no aggregate, or fewer than 10% fewer rows)."""
        L = nat.lib()
        n = A.shape[0]
        dev = A.device
        vals = A.values
        if vals.dtype != torch.float64:
            v64 = torch.empty(vals.shape[0], dtype=torch.float64, device=dev)
            nat.check(L.fa_cast(nat.dtype_code(vals.dtype), nat.FA_F64, vals.shape[0], nat.ptr(vals), nat.ptr(v64),
                                nat.stream()), "fa_cast")
            vals = v64
        agg = torch.empty(n, dtype=torch.int32, device=dev)
        ws = nat.workspace(L.fa_amg_aggregate_work_bytes(n, A.nnz), dev)
        nagg, rounds = ctypes.c_int64(0), ctypes.c_int32(0)
        nat.check(L.fa_amg_aggregate(nat.FA_F64, n, nat.ptr(A.crow), nat.ptr(A.col), nat.ptr(vals), float(self.theta),
                                     nat.ptr(agg), ctypes.byref(nagg), ctypes.byref(rounds), nat.ptr(ws), ws.numel(),
                                     nat.stream()), "fa_amg_aggregate")
        nc = int(nagg.value)
        if nc == 0 or nc > 0.9 * n:
            return None
        self.aggregation_rounds.append(int(rounds.value))
        keep = agg >= 0
        rows = torch.arange(n, device=dev)[keep]
        P0 = _csr(rows, agg[keep].long(), torch.ones(rows.shape[0], dtype=torch.float64, device=dev), n, nc)
        if A.nnz > self.smooth_max_row * n:
            return P0
        # A P0 on the Galerkin kernels (R = I): rows of nodes outside every
        # aggregate come out as the symbolic kernel's unit placeholder and are
        # dropped below
        Af = type(A)(A.crow, A.col, vals, A.shape)
        ident = _identity(n, dev)
        ap, ap_max = self._rap_pattern(ident, Af, P0, nc, max(_max_row(Af), 1))
        nat.check(L.fa_rap_values(nat.FA_F64, n, nat.ptr(ident.crow), nat.ptr(ident.col), nat.ptr(ident.values),
                                  nat.ptr(Af.crow), nat.ptr(Af.col), nat.ptr(vals), nat.ptr(P0.crow), nat.ptr(P0.col),
                                  nat.ptr(P0.values), nat.ptr(ap.crow), nat.ptr(ap.col), ap_max, nat.ptr(ap.values),
                                  nat.stream()), "fa_rap_values")
        dinv = torch.empty(n, dtype=torch.float64, device=dev)
        bound = ctypes.c_double(0.0)
        wsr = nat.workspace(L.fa_reduce_work_bytes(), dev)
        nat.check(L.fa_jacobi_setup(nat.FA_F64, n, A.nnz, nat.ptr(A.crow), nat.ptr(A.col), nat.ptr(vals),
                                    nat.ptr(dinv), ctypes.byref(bound), nat.ptr(wsr), wsr.numel(), nat.stream()),
                  "fa_jacobi_setup")
        omega = self.omega / max(float(bound.value), 1e-300)
        pv = torch.empty(ap.nnz, dtype=torch.float64, device=dev)
        nat.check(L.fa_amg_smooth_prolongation(nat.FA_F64, n, nat.ptr(A.crow), nat.ptr(A.col), nat.ptr(vals),
                                               nat.ptr(agg), nat.ptr(ap.crow), nat.ptr(ap.col), nat.ptr(ap.values),
                                               omega, nat.ptr(pv), nat.stream()), "fa_amg_smooth_prolongation")
        prow = ap.row_index()
        keep = agg[prow] >= 0
        return _csr(prow[keep], ap.col[keep].long(), pv[keep], n, nc)
