"""Geometric multigrid preconditioner on nested ``from_box`` meshes.

Role of the reference's ``GAMGSolver`` (fealpy/solver/gamg_solver.py:66-301):
a hierarchy of operators A_l, transfers P_l / R_l = P_l^T, V-cycles used as
the preconditioner of CG.  Designed for the GPU:

* for a degree-p space the first coarse level is P1 on the same mesh
  (interpolation of P1 into Pp, exact; the p-transfer of
  lagrange_fe_space.py:237-258); vector spaces apply every transfer per
  component;
* geometric levels are the ``from_box`` meshes with every cell count halved
  (odd counts rounded up); P1 prolongation evaluates the coarse P1 function
  at the fine nodes in closed form (Freudenthal weights of the Kuhn /
  diagonal-split cells).  Even counts give nested grids, where a fine node
  with grid index g takes 1/2 of the coarse nodes floor(g/2) and ceil(g/2)
  (one node with weight 1 when g is even);
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
from ..sparse.sell import FaSell, FaSell16, SellMatrix
from ..sparse.tensors import CSRTensor
from .cg import maxit_arg


class FaMgLevel(ctypes.Structure):
    _fields_ = [("A", FaSell), ("dinv", ctypes.c_void_p), ("R", FaSell), ("P", FaSell), ("lmax", ctypes.c_double),
                ("b", ctypes.c_void_p), ("x", ctypes.c_void_p), ("x2", ctypes.c_void_p), ("d", ctypes.c_void_p),
                ("r", ctypes.c_void_p), ("Ah", FaSell16), ("diag", ctypes.c_void_p), ("width", ctypes.c_int32),
                ("grid", ctypes.c_int32 * 3), ("bd_fine", ctypes.c_void_p), ("bd_coarse", ctypes.c_void_p)]


# nested 3D P1 grid levels apply R and P from the grid (k_box_restrict,
# k_box_prolong) instead of their SELL matrices: the same values
GRID_TRANSFERS = True


class FaMg(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int), ("nlevels", ctypes.c_int), ("levels", ctypes.POINTER(FaMgLevel)),
                ("degree", ctypes.c_int), ("lower", ctypes.c_double), ("upper", ctypes.c_double),
                ("coarse_chol", ctypes.c_void_p), ("coarse_degree", ctypes.c_int)]


def _grid_shape(mesh):
    meta = mesh.meshdata.get('from_box')
    if meta is None:
        raise ValueError("GeometricMG needs a full from_box mesh (nested halving is closed form only there); "
                         "AlgebraicMG builds its hierarchy from the matrix of any mesh")
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


def _coarse_dims(dims):
    """Cell counts of the next coarser grid: every count >= 2 halved, rounded up."""
    return [(n + 1) // 2 if n >= 2 else n for n in dims]


def _grid_prolongation(dims_f, device):
    """Scalar P1 prolongation coarse -> fine as (rows, cols, vals): each fine
    node takes the value of the coarse P1 function at its position.

    Both grids triangulate the same box with the diagonal-split (2D) / Kuhn
    (3D) cells of ``from_box``, whose simplices all contain the cube diagonal
    from the low corner to the high corner.  A point with local coordinates
    t in the coarse cube at corner c therefore lies in the simplex of the
    path c -> c + e_s1 -> c + e_s1 + e_s2 -> ... ordered by t_s1 >= t_s2 >= ...,
    and its barycentric weights are 1 - t_s1, t_s1 - t_s2, ..., t_sd (zero
    weights dropped).  An even count n (coarse n/2, nested grids) gives
    t in {0, 1/2}: the fine node takes coarse nodes floor(g/2) and ceil(g/2)
    with 1/2 each.  An odd count n (coarse (n+1)/2, not nested) gives
    fractional weights."""
    dims_c = _coarse_dims(dims_f)
    d = len(dims_f)
    nf = 1
    for n in dims_f:
        nf *= n + 1
    fine = torch.arange(nf, device=device)
    g = _grid_index(dims_f, fine)
    nn = torch.tensor(dims_f, dtype=torch.int64, device=device)
    mm = torch.tensor(dims_c, dtype=torch.int64, device=device)
    num = g * mm                                       # coarse coordinate = num / n, exactly
    base = num // nn
    t = (num % nn).to(torch.float64) / nn.to(torch.float64)
    ts, order = torch.sort(t, dim=-1, descending=True, stable=True)
    ext = torch.cat([torch.ones_like(ts[:, :1]), ts, torch.zeros_like(ts[:, :1])], dim=1)
    w = ext[:, :-1] - ext[:, 1:]                       # (nf, d+1): 1 - t_s1, t_s1 - t_s2, ..., t_sd
    step = torch.zeros(nf, d + 1, d, dtype=torch.int64, device=device)
    step[:, 1:, :] = torch.nn.functional.one_hot(order, d).cumsum(dim=1)
    vert = base[:, None, :] + step                     # (nf, d+1, d) path vertices
    cols = _node_ids(dims_c, vert)
    rows = fine[:, None].expand(nf, d + 1)
    keep = w != 0
    return rows[keep], cols[keep], w[keep]


def _coarse_to_fine_node(dims_f, dims_c, g):
    """Fine grid index nearest to coarse grid index g (2g on nested grids;
    boundary nodes map to boundary nodes)."""
    nn = torch.tensor(dims_f, dtype=torch.int64, device=g.device)
    mm = torch.tensor(dims_c, dtype=torch.int64, device=g.device)
    return (g * nn + mm // 2) // mm


def _order_prolongation(space):
    """Scalar P1 -> Pp interpolation on one mesh as (rows, cols, vals): the
    Pp DOF at barycentric point alpha/p of a cell takes alpha_v/p of the cell's
    vertex v (P1 is contained in Pp, so the transfer is exact).  Each DOF's row
    comes from its first cell (every cell gives the same row)."""
    mesh = space.mesh
    c2d = space.cell_to_dof().long()
    cell = mesh.cell.long()
    nc, ldof = c2d.shape
    gdof = space.number_of_global_dofs()
    dev = c2d.device
    first = torch.full((gdof,), nc * ldof, dtype=torch.int64, device=dev)
    first.scatter_reduce_(0, c2d.reshape(-1), torch.arange(nc * ldof, device=dev), reduce='amin')
    rc, ri = first // ldof, first % ldof
    mi = torch.as_tensor(space.multi_index_matrix(), dtype=torch.float64, device=dev)     # (ldof, TD+1)
    w = mi[ri] / space.p                                                                  # (gdof, TD+1)
    rows = torch.arange(gdof, device=dev)[:, None].expand_as(w).reshape(-1)
    cols = cell[rc].reshape(-1)
    vals = w.reshape(-1)
    keep = vals != 0
    return rows[keep], cols[keep], vals[keep]


def _block_prolongation(rows, cols, vals, k, prio, nf, nc, bd_f, bd_c):
    """Expand a scalar transfer to k components (component-interleaved, or
    component blocks when prio) and drop fine Dirichlet rows / coarse
    Dirichlet columns; -> CSR sorted by (row, col)."""
    if k > 1:
        comp = torch.arange(k, device=rows.device)
        if prio:
            rows = (comp[:, None] * nf + rows[None, :]).reshape(-1)
            cols = (comp[:, None] * nc + cols[None, :]).reshape(-1)
        else:
            rows = (rows[None, :] * k + comp[:, None]).reshape(-1)
            cols = (cols[None, :] * k + comp[:, None]).reshape(-1)
        vals = vals[None, :].expand(k, -1).reshape(-1)
    keep = ~bd_f[rows] & ~bd_c[cols]
    rows, cols, vals = rows[keep], cols[keep], vals[keep]
    order = torch.argsort(rows * (k * nc) + cols)
    return _csr(rows[order], cols[order], vals[order], k * nf, k * nc)


def _select_nodes(bd, k, prio, n_scalar, ids):
    """Flags of the scalar nodes `ids` for every component."""
    if k == 1:
        return bd[ids]
    b = bd.reshape(k, n_scalar) if prio else bd.reshape(n_scalar, k).T
    sel = b[:, ids]
    return sel.reshape(-1) if prio else sel.T.reshape(-1)


def _csr(rows, cols, vals, nrows, ncols):
    crow = torch.zeros(nrows + 1, dtype=torch.int64, device=rows.device)
    crow[1:] = torch.cumsum(torch.bincount(rows, minlength=nrows), 0)
    return CSRTensor(crow.to(torch.int32), cols.to(torch.int32), vals, (nrows, ncols))


def _transpose(M):
    rows = M.row_index()
    cols = M.col.long()
    order = torch.argsort(cols * M.shape[0] + rows)
    return _csr(cols[order], rows[order], M.values[order], M.shape[1], M.shape[0])


def _identity(n, device):
    idx = torch.arange(n, dtype=torch.int32, device=device)
    crow = torch.arange(n + 1, dtype=torch.int32, device=device)
    return CSRTensor(crow, idx, torch.ones(n, dtype=torch.float64, device=device), (n, n))


def _max_row(M):
    return int((M.crow[1:] - M.crow[:-1]).max().item()) if M.shape[0] else 0


# candidates per row the symbolic Galerkin kernels hold (fa_rap_count's largest LDS list)
RAP_CANDIDATES = 8192
# Galerkin products whose coarse rows' fine rows hold at most this many
# entries use the staged term words (fa_rap_values_staged); False keeps the
# 32-bit words with A entry indices (fa_rap_values_terms), the same values
RAP_STAGED = True
RAP_STAGE_MAX = 512
RAP_ROW_WORDS = 1024


def _row_sums(M, w):
    """sum over each row of M of w[col] (int64); a column outside w counts 0
    (the unit placeholder the symbolic kernel gives a row without
    candidates, whose column is the row index, which R never reaches)."""
    col = M.col.long()
    inside = col < w.shape[0]
    v = torch.where(inside, w[torch.clamp(col, max=max(w.shape[0] - 1, 0))], torch.zeros_like(col))
    cs = torch.zeros(v.shape[0] + 1, dtype=torch.int64, device=v.device)
    cs[1:] = torch.cumsum(v, 0)
    crow = M.crow.long()
    return cs[crow[1:]] - cs[crow[:-1]]


def _stage_fits(R, A, C, tptr):
    """Whether every coarse row of R A P is within the staged form's bounds:
    at most 64 R entries and 63 coarse entries, at most RAP_ROW_WORDS terms,
    and at most RAP_STAGE_MAX entries in the A rows R row I names."""
    if R.nnz == 0 or C.nnz == 0:
        return False
    per_r = _row_sums(R, (A.crow[1:] - A.crow[:-1]).long())
    crow = C.crow.long()
    words = tptr.long()[crow[1:]] - tptr.long()[crow[:-1]]
    return (int(per_r.max()) <= RAP_STAGE_MAX and int((R.crow[1:] - R.crow[:-1]).max()) <= 64
            and int((C.crow[1:] - C.crow[:-1]).max()) <= 63 and int(words.max()) <= RAP_ROW_WORDS)


def _candidates(R, A, P):
    """Largest number of (R entry, A entry, P entry) triples of one row of
    R A P: the exact length of the symbolic kernel's candidate list (at
    least 1)."""
    plen = (P.crow[1:] - P.crow[:-1]).long()
    per_a = _row_sums(A, plen)
    per_r = _row_sums(R, per_a)
    return max(int(per_r.max().item()) if per_r.numel() else 1, 1)


class _Level:
    pass


class GeometricMG:
    """V-cycle preconditioner.  ``space`` is a Lagrange space (any p) or a
    TensorFunctionSpace over one, on a ``from_box`` mesh; ``is_bd`` flags the
    Dirichlet DOFs of that space.  For p > 1 the first transfer goes to P1 on
    the same mesh; vector spaces coarsen every component alike."""

    def __init__(self, space, is_bd, *, dtype=torch.float32, degree=2, lower=0.1, upper=1.0, coarse_max=64,
                 power_iters=12, coarse_degree=12, compress=True, graphs=False):
        ss = getattr(space, 'scalar_space', space)
        k = int(getattr(space, 'dof_numel', 1))
        prio = bool(getattr(space, 'dof_priority', False))
        if getattr(space, 'dof_ndim', 1) != 1:
            raise NotImplementedError("GeometricMG covers scalar and vector spaces")
        mesh = ss.mesh
        dims = _grid_shape(mesh)
        dev = space.device
        self.dtype, self.degree, self.lower, self.upper = dtype, degree, lower, upper
        # FP32 cycles smooth with FP16 D^-1 A and int16 column offsets where the band allows
        self.compress = bool(compress) and dtype == torch.float32
        self.power_iters = power_iters
        # graphs: PCG iterations replayed from a HIP graph (fa_pcg_graph), for
        # launch-bound solves; the handle is made on first use
        self.graphs = bool(graphs)
        self._graph = None
        self.device = dev
        self.levels = []
        bd = is_bd.to(torch.bool).contiguous()
        if bd.shape[0] != space.number_of_global_dofs():
            raise ValueError("is_bd must flag every DOF of the space")
        lev = _Level()
        lev.n, lev.bd = space.number_of_global_dofs(), bd
        self.levels.append(lev)
        if ss.p > 1:                                    # Pp -> P1 on the same mesh
            nn = mesh.number_of_nodes()
            ns = ss.number_of_global_dofs()
            bd_c = _select_nodes(bd, k, prio, ns, torch.arange(nn, device=dev))
            self._link(lev, _block_prolongation(*_order_prolongation(ss), k, prio, ns, nn, bd, bd_c))
            nxt = _Level()
            nxt.n, nxt.bd = k * nn, bd_c
            self.levels.append(nxt)
            lev, bd = nxt, bd_c
        lev.dims = dims
        while any(n >= 2 for n in dims) and lev.n > coarse_max:
            dims_c = _coarse_dims(dims)
            nf_s, nc_s = 1, 1
            for n in dims:
                nf_s *= n + 1
            for n in dims_c:
                nc_s *= n + 1
            g = _grid_index(dims_c, torch.arange(nc_s, device=dev))
            bd_c = _select_nodes(bd, k, prio, nf_s, _node_ids(dims, _coarse_to_fine_node(dims, dims_c, g)))
            self._link(lev, _block_prolongation(*_grid_prolongation(dims, dev), k, prio, nf_s, nc_s, bd, bd_c))
            # nested scalar 3D grids: the transfers are the Kuhn-edge stencils
            lev.grid_transfer = tuple(dims) if (k == 1 and len(dims) == 3 and all(n % 2 == 0 for n in dims)) else None
            nxt = _Level()
            nxt.dims, nxt.n, nxt.bd = dims_c, k * nc_s, bd_c
            self.levels.append(nxt)
            lev, dims, bd = nxt, dims_c, bd_c
        self.coarse_degree = coarse_degree
        self._finish_levels()
        self._pattern_key = None
        self._desc = None
        self._setup_A = None

    def _finish_levels(self):
        """Work vectors of every level; the coarsest level is solved exactly
        (dense inverse) up to 2048 rows, above that (a coarse_max above 2048)
        smoothed from zero with a Chebyshev polynomial of degree coarse_degree."""
        self.coarse_dense = self.levels[-1].n <= 2048
        for lv in self.levels:
            z = lambda: torch.zeros(lv.n, dtype=self.dtype, device=self.device)
            lv.b, lv.x, lv.x2, lv.d, lv.r, lv.dinv = z(), z(), z(), z(), z(), z()

    def _link(self, lev, P):
        lev.P = P
        lev.R = _transpose(P)
        lev.Ps = SellMatrix(P.crow, P.col, P.shape)
        lev.Rs = SellMatrix(lev.R.crow, lev.R.col, lev.R.shape)
        lev.Ps.set_values(P.values, self.dtype)
        lev.Rs.set_values(lev.R.values, self.dtype)

    # -- per-matrix setup -----------------------------------------------------
    def _symbolic(self, A):
        from ..sparse.sell import sell_of
        self.levels[0].A = A
        self.levels[0].As = sell_of(A)
        self.levels[0].max_row = _max_row(A)
        for lv in self.levels:
            lv.compressed = None        # decided by the next setup (band check of the new pattern)
        for l in range(len(self.levels) - 1):
            self._symbolic_pair(l)
        last = self.levels[-1]
        last.dense = (torch.zeros((last.n, last.n), dtype=torch.float64, device=self.device) if self.coarse_dense
                      else None)

    def _symbolic_pair(self, l):
        """Pattern of the Galerkin product of levels l -> l+1 and its stages."""
        f, c = self.levels[l], self.levels[l + 1]
        cand = _candidates(f.R, f.A, f.P)
        # stages name the level's own operator None: the numeric setup reads
        # the current f.A (a new matrix on the same pattern keeps the stages)
        if cand <= RAP_CANDIDATES:
            c.A, c.max_row = self._rap_pattern(f.R, f.A, f.P, c.n, cand)
            f.stages = [(f.R, None, f.P, c.A, c.max_row, self._rap_terms(f.R, f.A, f.P, c.A, cand))]
        else:
            # R (A P) in two products: each row's candidate list stays short
            # (P3 -> P1 in 3D: 975 against 34,125 for R A P at once)
            i_f, i_c = _identity(f.n, self.device), _identity(c.n, self.device)
            c1 = _candidates(i_f, f.A, f.P)
            ap, ap_max = self._rap_pattern(i_f, f.A, f.P, c.n, c1)
            c2 = _candidates(f.R, ap, i_c)
            c.A, c.max_row = self._rap_pattern(f.R, ap, i_c, c.n, c2)
            f.stages = [(i_f, None, f.P, ap, ap_max, self._rap_terms(i_f, f.A, f.P, ap, c1)),
                        (f.R, ap, i_c, c.A, c.max_row, self._rap_terms(f.R, ap, i_c, c.A, c2))]
        c.As = SellMatrix(c.A.crow, c.A.col, c.A.shape)

    def _numeric_pair(self, l, vd):
        """Values of the Galerkin product of levels l -> l+1 (in type vd)."""
        L = nat.lib()
        vcode = nat.dtype_code(vd)
        for R, Af, P, C, cmax, terms in self.levels[l].stages:
            Af = self.levels[l].A if Af is None else Af
            if C.values.dtype != vd:            # value arrays follow the operator's type
                C._values = torch.zeros(C.nnz, dtype=vd, device=self.device)
            if terms is not None:
                tptr, tw, wset, nw, staged = terms
                if staged:
                    nat.check(L.fa_rap_values_staged(vcode, C.shape[0], nat.ptr(R.crow), nat.ptr(R.col),
                                                     nat.ptr(Af.crow), nat.ptr(C.crow), nat.ptr(tptr), nat.ptr(tw),
                                                     wset, nw, nat.ptr(Af.values), nat.ptr(C.values),
                                                     nat.stream()), "fa_rap_values_staged")
                else:
                    nat.check(L.fa_rap_values_terms(vcode, C.nnz, nat.ptr(tptr), nat.ptr(tw), wset, nw,
                                                    nat.ptr(Af.values), nat.ptr(C.values), nat.stream()),
                              "fa_rap_values_terms")
            else:
                nat.check(L.fa_rap_values(vcode, C.shape[0], nat.ptr(R.crow), nat.ptr(R.col), nat.ptr(R.values),
                                          nat.ptr(Af.crow), nat.ptr(Af.col), nat.ptr(Af.values),
                                          nat.ptr(P.crow), nat.ptr(P.col), nat.ptr(P.values), nat.ptr(C.crow),
                                          nat.ptr(C.col), cmax, nat.ptr(C.values), nat.stream()),
                          "fa_rap_values")

    def _rap_pattern(self, R, A, P, ncols, cand):
        """Symbolic product R A P (fa_rap_count / fa_rap_fill) -> (CSR with
        zero values, longest row).  A row with no candidates gets a unit
        diagonal; in the first product of two (A P, rows of fine DOFs) that
        only happens for Dirichlet rows, which R never reads."""
        L = nat.lib()
        nrows = R.shape[0]
        # the kernel compacts a full candidate list to its distinct columns,
        # so more candidates than its capacity still fit when the distinct
        # columns (plus one batch) do; it reports FA_ECAPACITY otherwise
        cand = min(max(int(cand), 1), RAP_CANDIDATES)
        cptr = torch.empty(nrows + 1, dtype=torch.int32, device=self.device)
        ws = nat.workspace(L.fa_rap_work_bytes(nrows), self.device)
        nnz, mrow = ctypes.c_int64(0), ctypes.c_int32(0)
        nat.check(L.fa_rap_count(nrows, ncols, nat.ptr(R.crow), nat.ptr(R.col), nat.ptr(A.crow), nat.ptr(A.col),
                                 nat.ptr(P.crow), nat.ptr(P.col), cand, nat.ptr(cptr), ctypes.byref(nnz),
                                 ctypes.byref(mrow), nat.ptr(ws), ws.numel(), nat.stream()), "fa_rap_count")
        cidx = torch.empty(max(int(nnz.value), 1), dtype=torch.int32, device=self.device)
        nat.check(L.fa_rap_fill(nrows, nat.ptr(R.crow), nat.ptr(R.col), nat.ptr(A.crow), nat.ptr(A.col),
                                nat.ptr(P.crow), nat.ptr(P.col), cand, nat.ptr(cptr), nat.ptr(cidx), nat.stream()),
                  "fa_rap_fill")
        C = CSRTensor(cptr, cidx[:int(nnz.value)], torch.zeros(int(nnz.value), dtype=torch.float64,
                                                              device=self.device), (nrows, ncols))
        return C, int(mrow.value)

    def _rap_terms(self, R, A, P, C, cand):
        """Term list of the numeric product (fa_rap_terms_*), or None when the
        transfers are empty or take more than 3 distinct values, or A is too
        large; the search kernel (fa_rap_values) then forms the product."""
        L = nat.lib()
        uniq = torch.unique(torch.cat([R.values, P.values]))
        if uniq.numel() == 0 or uniq.numel() > 3 or A.nnz >= (1 << 28) or C.nnz == 0:
            return None
        vals = uniq.cpu().tolist()
        wset = (ctypes.c_double * len(vals))(*vals)
        ne = C.nnz
        tptr = torch.empty(ne + 1, dtype=torch.int32, device=self.device)
        ws = nat.workspace(L.fa_rap_terms_work_bytes(ne), self.device)
        nterms = ctypes.c_int64(0)
        args = (C.shape[0], nat.ptr(R.crow), nat.ptr(R.col), nat.ptr(R.values), nat.ptr(A.crow), nat.ptr(A.col),
                nat.ptr(P.crow), nat.ptr(P.col), nat.ptr(P.values), max(int(cand), 1), nat.ptr(C.crow), ne, wset,
                len(vals))
        nat.check(L.fa_rap_terms_count(*args, nat.ptr(tptr), ctypes.byref(nterms), nat.ptr(ws), ws.numel(),
                                       nat.stream()), "fa_rap_terms_count")
        staged = RAP_STAGED and _stage_fits(R, A, C, tptr)
        if staged:   # 16-bit words naming the A entry's slot in the coarse row's staged fine rows
            terms = torch.empty(max(int(nterms.value), 1), dtype=torch.int16, device=self.device)
            nat.check(L.fa_rap_terms_fill_staged(*args, nat.ptr(tptr), int(nterms.value), nat.ptr(terms),
                                                 nat.ptr(ws), ws.numel(), nat.stream()), "fa_rap_terms_fill_staged")
        else:
            terms = torch.empty(max(int(nterms.value), 1), dtype=torch.int32, device=self.device)
            nat.check(L.fa_rap_terms_fill(*args, nat.ptr(tptr), int(nterms.value), nat.ptr(terms), nat.ptr(ws),
                                          ws.numel(), nat.stream()), "fa_rap_terms_fill")
        return (tptr, terms, wset, len(vals), staged)

    def setup(self, A):
        """Numeric setup for the (Dirichlet-applied) fine matrix A, FP64 or
        FP32.  The coarse operators keep A's value type (an operator assembled
        in FP32 gets FP32 Galerkin products, each formed in FP64 and rounded
        once); the cycle runs in self.dtype either way."""
        if A.dtype not in (torch.float64, torch.float32):
            raise TypeError("GeometricMG.setup expects an FP64 or FP32 matrix")
        key = (A.crow.data_ptr(), A.col.data_ptr(), A.nnz)
        if key != self._pattern_key:
            self._symbolic(A)
            self._pattern_key = key
        L = nat.lib()
        vd = A.dtype
        vcode = nat.dtype_code(vd)
        self.levels[0].A = A
        for l in range(len(self.levels) - 1):
            self._numeric_pair(l, vd)
        code = nat.dtype_code(self.dtype)
        ws = nat.workspace(L.fa_reduce_work_bytes(), self.device)
        bounds = torch.empty(len(self.levels), dtype=torch.float64, device=self.device)
        f32 = self.dtype == torch.float32
        for i, lv in enumerate(self.levels):
            lv.As.set_values(lv.A.values, vd)       # SELL copy in A's type (level 0: the outer operator)
            if getattr(lv, 'compressed', None) is None:            # first setup of this pattern: band check
                lv.compressed = self.compress and lv.n > 1024 and lv.As.set_compressed(lv.A.values, ws)
                lv.diag = torch.empty_like(lv.dinv) if lv.compressed else None
            v32 = None
            if f32 and not lv.compressed and vd == torch.float64:
                v32 = lv.As.value_buffer(torch.float32, self.device)
            if not f32 and vd == torch.float32:     # an FP64 cycle over an FP32 operator: its FP64 copy
                lv.As.set_values(lv.A.values, torch.float64)
            din = lv.As.desc(vd)
            h16 = lv.As.desc16() if lv.compressed else None
            nat.check(L.fa_smoother_setup(code, vcode, ctypes.byref(din), nat.ptr(lv.dinv),
                                          nat.ptr(lv.diag) if lv.compressed else None,
                                          nat.ptr(v32) if v32 is not None else None,
                                          ctypes.byref(h16) if h16 is not None else None,
                                          bounds[i:].data_ptr(), nat.ptr(ws), ws.numel(), nat.stream()),
                      "fa_smoother_setup")
        # the coarsest inverse is queued before the host waits for the bounds
        last = self.levels[-1]
        if self.coarse_dense:
            cv = last.A.values
            if cv.dtype != torch.float64:
                c64 = torch.empty(cv.shape[0], dtype=torch.float64, device=self.device)
                nat.check(L.fa_cast(vcode, nat.FA_F64, cv.shape[0], nat.ptr(cv), nat.ptr(c64), nat.stream()),
                          "fa_cast")
                cv = c64
            nat.check(L.fa_csr_to_dense(last.n, nat.ptr(last.A.crow), nat.ptr(last.A.col), nat.ptr(cv),
                                        nat.ptr(last.dense), nat.stream()), "fa_csr_to_dense")
            nat.check(L.fa_dense_spd_inverse(last.n, nat.ptr(last.dense), nat.stream()), "fa_dense_spd_inverse")
        for lv, lam in zip(self.levels, bounds.tolist()):           # one copy for every level's bound
            lv.lmax = float(lam)
        self._build_desc()
        self._setup_A = A
        self._setup_version = A.values._version
        return self

    def estimate_lmax(self, level, iters=None):
        """Power-iteration estimate of lambda_max(D^-1 A) on a level (diagnostics)."""
        L = nat.lib()
        lv = self.levels[level]
        lv.As.set_values(lv.A.values, self.dtype)     # compressed levels keep no current copy in the cycle dtype
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
            lv.As.value_buffer(self.dtype, self.device)
            e.A = lv.As.desc(self.dtype)
            e.dinv = lv.dinv.data_ptr()
            e.R = lv.Rs.desc(self.dtype) if i < n - 1 else empty
            e.P = lv.Ps.desc(self.dtype) if i < n - 1 else empty
            e.lmax = getattr(lv, 'lmax', 2.0)
            e.width = getattr(lv, 'max_row', 0)
            e.b, e.x, e.x2, e.d, e.r = (lv.b.data_ptr(), lv.x.data_ptr(), lv.x2.data_ptr(), lv.d.data_ptr(),
                                        lv.r.data_ptr())
            if getattr(lv, 'compressed', False):
                e.Ah = lv.As.desc16()
                e.diag = lv.diag.data_ptr()
            else:
                e.Ah = FaSell16(0, None, None, None)
                e.diag = None
            gt = getattr(lv, 'grid_transfer', None) if (GRID_TRANSFERS and i < n - 1) else None
            if gt is not None:     # Dirichlet flags as bytes (torch.bool is one byte a flag)
                e.grid = (ctypes.c_int32 * 3)(*gt)
                e.bd_fine = lv.bd.data_ptr()
                e.bd_coarse = self.levels[i + 1].bd.data_ptr()
            else:
                e.grid = (ctypes.c_int32 * 3)(0, 0, 0)
                e.bd_fine = e.bd_coarse = None
        self._lv_array = arr
        dense = self.levels[-1].dense
        self._desc = FaMg(nat.dtype_code(self.dtype), n, arr, self.degree, self.lower, self.upper,
                          dense.data_ptr() if dense is not None else None, self.coarse_degree)

    def __del__(self):
        g = getattr(self, '_graph', None)
        if g is not None and nat._lib is not None:
            nat._lib.fa_pcg_graph_destroy(g)
            self._graph = None

    # -- use ----------------------------------------------------------------------
    def __matmul__(self, r):
        """z = V(r): one V-cycle of the hierarchy set up last (the cycle runs in
        self.dtype; z has r's dtype)."""
        if self._desc is None:
            raise RuntimeError("GeometricMG: call setup(A) before applying the cycle")
        rc = r.to(self.dtype).contiguous()
        z = torch.empty_like(rc)
        nat.check(nat.lib().fa_mg_vcycle(ctypes.byref(self._desc), nat.ptr(rc), nat.ptr(z), nat.stream()),
                  "fa_mg_vcycle")
        return z if z.dtype == r.dtype else z.to(r.dtype)

    def ensure_setup(self, A):
        """Set up for A unless the last setup was for this matrix and its values
        have not changed since."""
        fresh = self._setup_A is A and getattr(self, '_setup_version', None) == A.values._version
        if not fresh:
            self.setup(A)
            self._setup_A = A
            self._setup_version = A.values._version

    def pcg(self, A, b, x0=None, *, atol=1e-12, rtol=1e-8, maxit=10000, check_every=4):
        """Preconditioned CG on the device: A's dtype outside, V-cycle in self.dtype."""
        from ..sparse.sell import sell_of
        self.ensure_setup(A)
        L = nat.lib()
        dt = A.dtype
        if self.levels[0].A is A:
            S = self.levels[0].As          # its copy in A's type was just filled from A by setup()
        else:
            S = sell_of(A)
        b = b.to(dt).contiguous()
        # from zero (no x0, no graph): fa_pcg_from_zero writes x = 0 itself and
        # skips the product A x0
        from_zero = x0 is None and not self.graphs
        x = torch.empty_like(b) if from_zero else (torch.zeros_like(b) if x0 is None else x0.to(dt).clone().contiguous())
        ws = nat.workspace(L.fa_pcg_work_bytes(nat.dtype_code(dt), b.shape[0]), b.device)
        niter, res = ctypes.c_int64(0), ctypes.c_double(0.0)
        d = S.desc(dt)
        args = (nat.dtype_code(dt), ctypes.byref(d), ctypes.byref(self._desc), nat.ptr(b), nat.ptr(x), float(atol),
                float(rtol), maxit_arg(maxit), int(check_every), ctypes.byref(niter), ctypes.byref(res), nat.ptr(ws),
                ws.numel())
        if self.graphs:
            if self._graph is None:
                h = ctypes.c_void_p()
                nat.check(L.fa_pcg_graph_create(ctypes.byref(h)), "fa_pcg_graph_create")
                self._graph = h
            st = L.fa_pcg_graph(*args, self._graph, nat.stream())
        else:
            st = (L.fa_pcg_from_zero if from_zero else L.fa_pcg)(*args, nat.stream())
        if st != nat.FA_ENOCONV:
            nat.check(st, "fa_pcg")
        return x, {'residual': float(res.value), 'niter': int(niter.value)}


# ---- many small solves in one launch: one workgroup per system -------------------
class FaWgLevel(ctypes.Structure):
    _fields_ = [("A", FaSell), ("R", FaSell), ("P", FaSell), ("dinv", ctypes.c_void_p), ("b", ctypes.c_void_p),
                ("x", ctypes.c_void_p), ("x2", ctypes.c_void_p), ("d", ctypes.c_void_p), ("r", ctypes.c_void_p),
                ("lmax", ctypes.c_double)]


WG_MAX_LEVELS = 16


class FaWgSolve(ctypes.Structure):
    _fields_ = [("A", FaSell), ("b", ctypes.c_void_p), ("x", ctypes.c_void_p), ("r", ctypes.c_void_p),
                ("p", ctypes.c_void_p), ("Ap", ctypes.c_void_p), ("atol", ctypes.c_double), ("rtol", ctypes.c_double),
                ("maxit", ctypes.c_int64), ("nlevels", ctypes.c_int32), ("degree", ctypes.c_int32),
                ("lower", ctypes.c_double), ("upper", ctypes.c_double), ("coarse_inv", ctypes.c_void_p),
                ("levels", FaWgLevel * WG_MAX_LEVELS), ("niter", ctypes.c_int64), ("residual", ctypes.c_double),
                ("status", ctypes.c_int32), ("pad", ctypes.c_int32)]


def pcg_workgroups(systems, *, atol=0.0, rtol=1e-8, maxit=10000):
    """Solve several small FP64 systems at once (fa_pcg_wg_batched), one
    1024-thread workgroup per system running its whole MG-preconditioned CG:
    the FP64 recurrence on A and the V-cycle of its hierarchy, with block
    barriers between phases and no kernel launch per step.

    systems: [(mg, A, b)], mg a GeometricMG with an FP32 cycle built with
    compress=False and set up for A (FP64, Dirichlet-applied), b FP64.
    Returns [(x, {'niter', 'residual', 'status'})] in order; each result
    depends only on its own system (the same alone and in any batch)."""
    L = nat.lib()
    n = len(systems)
    if n == 0:
        return []
    reqs = (FaWgSolve * n)()
    keep, outs = [], []
    for q, (mg, A, b) in enumerate(systems):
        mg.ensure_setup(A)
        if mg.dtype != torch.float32 or any(getattr(lv, 'compressed', False) for lv in mg.levels):
            raise ValueError("pcg_workgroups needs an FP32 cycle built with compress=False")
        if A.dtype != torch.float64 or mg.levels[0].A is not A:
            raise ValueError("pcg_workgroups needs the FP64 operator the hierarchy was set up for")
        if len(mg.levels) > WG_MAX_LEVELS or mg.levels[-1].dense is None:
            raise ValueError("pcg_workgroups needs at most 16 levels and a dense coarsest inverse")
        nrow = A.shape[0]
        dev = b.device
        bb = b.to(torch.float64).contiguous()
        x = torch.zeros(nrow, dtype=torch.float64, device=dev)
        work = torch.empty((3, nrow), dtype=torch.float64, device=dev)
        keep.append((bb, work))
        outs.append(x)
        e = reqs[q]
        e.A = mg.levels[0].As.desc(torch.float64)
        e.b, e.x = bb.data_ptr(), x.data_ptr()
        e.r, e.p, e.Ap = (work[i].data_ptr() for i in range(3))
        e.atol, e.rtol, e.maxit = float(atol), float(rtol), max(int(maxit), 1)
        e.nlevels, e.degree, e.lower, e.upper = len(mg.levels), mg.degree, mg.lower, mg.upper
        e.coarse_inv = mg.levels[-1].dense.data_ptr()
        for l, src in enumerate(mg._lv_array):
            w = e.levels[l]
            w.A, w.R, w.P, w.dinv = src.A, src.R, src.P, src.dinv
            w.b, w.x, w.x2, w.d, w.r, w.lmax = src.b, src.x, src.x2, src.d, src.r, src.lmax
    raw = bytes(memoryview(reqs).cast("B"))
    dev = systems[0][2].device
    dreq = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(dev)
    nat.check(L.fa_pcg_wg_batched(n, reqs, ctypes.c_void_p(dreq.data_ptr()), nat.stream()), "fa_pcg_wg_batched")
    back = dreq.cpu()                                   # waits for the stream
    done = (FaWgSolve * n).from_buffer_copy(back.numpy().tobytes())
    res = []
    for q in range(n):
        info = {'niter': int(done[q].niter), 'residual': float(done[q].residual), 'status': int(done[q].status)}
        res.append((outs[q], info))
    return res
