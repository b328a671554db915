"""Solver kernels on the GPU: SELL SpMV, Galerkin RAP, V-cycle, MG-PCG.

Oracles: SciPy (SpMV, triple products), the textbook identity that for nested
P1 spaces and a constant coefficient R A_h P equals the coarse assembled
matrix, symmetry of the V-cycle, and the tight FP64 CG solution."""
import numpy as np
import pytest
import torch

import fealpy_amd as fa
from fealpy_amd.solver import GeometricMG
from fealpy_amd.sparse.sell import SellMatrix

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")
PI = np.pi


def _system(n, k=None, p=1, dim=3):
    if dim == 3:
        mesh = fa.TetrahedronMesh.from_box(nx=n, ny=n, nz=n, device=DEV)
        u = fa.cartesian(lambda x: torch.sin(PI * x[..., 0]) * torch.sin(PI * x[..., 1]) * torch.sin(PI * x[..., 2]))
    else:
        mesh = fa.TriangleMesh.from_box(nx=n, ny=n, device=DEV)
        u = fa.cartesian(lambda x: torch.sin(PI * x[..., 0]) * torch.sin(PI * x[..., 1]))
    space = fa.LagrangeFESpace(mesh, p)
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=2)).assembly()
    b = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(1.0, q=2)).assembly()
    bc = fa.DirichletBC(space, gd=u)
    A2, b2 = bc.apply(A, b)
    return mesh, space, bc, A2, b2


def _lognormal(mesh, seed=5):
    g = np.random.default_rng(seed).standard_normal(mesh.number_of_cells())
    return torch.as_tensor(np.exp(g), device=DEV)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_sell_spmv_matches_scipy(dtype):
    mesh, space, bc, A, b = _system(10, p=2)
    S = SellMatrix(A.crow, A.col, A.shape)
    S.set_values(A.values, dtype)
    x = torch.rand(A.shape[0], dtype=torch.float64, device=DEV)
    y = S.matvec(x.to(dtype))
    ref = A.to_scipy() @ x.cpu().numpy()
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    np.testing.assert_allclose(y.double().cpu().numpy(), ref, rtol=0, atol=tol * np.abs(ref).max())
    assert S.entries >= A.nnz and S.entries % 64 == 0


def test_rap_equals_coarse_assembly():
    """Constant coefficient: Galerkin R A_h P == A_2h (nested P1 spaces), after BC."""
    n = 16
    mesh, space, bc, A, b = _system(n)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float64, coarse_max=1)
    mg.setup(A)
    for l in range(1, len(mg.levels)):
        m = n >> l
        _, _, _, Ac, _ = _system(m)
        C = mg.levels[l].A
        assert torch.equal(C.crow, Ac.crow) and torch.equal(C.col, Ac.col), f"pattern level {l}"
        scale = float(Ac.values.abs().max())
        assert float((C.values - Ac.values).abs().max()) < 1e-12 * scale, f"values level {l}"


def test_rap_matches_scipy_triple_product():
    mesh, space, bc, A, b = _system(8, k=None)
    k = _lognormal(mesh)
    _, _, _, A, _ = _system(8, k=k)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float64, coarse_max=1)
    mg.setup(A)
    f, c = mg.levels[0], mg.levels[1]
    ref = (f.R.to_scipy() @ f.A.to_scipy() @ f.P.to_scipy()).tocsr()
    got = c.A.to_scipy()
    empty = np.diff(ref.indptr) == 0
    ref = ref + __import__("scipy").sparse.diags(empty.astype(float))
    diff = abs(got - ref)
    assert diff.max() < 1e-12 * abs(ref).max()


def _rap_by_search(f, c):
    """The same product by the pattern-search kernel (fa_rap_values)."""
    import ctypes
    from fealpy_amd import _native as nat
    out = torch.empty_like(c.A.values)
    nat.check(nat.lib().fa_rap_values(nat.FA_F64, c.n, nat.ptr(f.R.crow), nat.ptr(f.R.col), nat.ptr(f.R.values),
                                      nat.ptr(f.A.crow), nat.ptr(f.A.col), nat.ptr(f.A.values), nat.ptr(f.P.crow),
                                      nat.ptr(f.P.col), nat.ptr(f.P.values), nat.ptr(c.A.crow), nat.ptr(c.A.col),
                                      c.max_row, nat.ptr(out), nat.stream()), "fa_rap_values")
    return out


@pytest.mark.parametrize("p", [1, 2])
def test_rap_term_list(p):
    """Term-list numeric product equals SciPy's triple product and the search
    kernel on every level (transfers 1, 1/2: nested P1 grids, P2 -> P1)."""
    import scipy.sparse as sp
    n = {1: 8, 2: 6}[p]
    mesh, space, bc, A, b = _system(n, p=p)
    k = _lognormal(mesh, seed=11)
    _, _, _, A, _ = _system(n, k=k, p=p)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float64, coarse_max=1)
    mg.setup(A)
    assert mg.levels[0].stages[0][-1] is not None      # term list
    for l in range(len(mg.levels) - 1):
        f, c = mg.levels[l], mg.levels[l + 1]
        ref = (f.R.to_scipy() @ f.A.to_scipy() @ f.P.to_scipy()).tocsr()
        ref = ref + sp.diags((np.diff(ref.indptr) == 0).astype(float))
        got = c.A.to_scipy()
        scale = abs(ref).max()
        assert abs(got - ref).max() < 1e-12 * scale, f"level {l}"
        other = _rap_by_search(f, c)
        assert float((other - c.A.values).abs().max()) < 1e-13 * scale, f"level {l}"


def _csr_dev(M):
    M = M.tocsr()
    M.sort_indices()
    return (torch.as_tensor(M.indptr, dtype=torch.int32, device=DEV),
            torch.as_tensor(M.indices, dtype=torch.int32, device=DEV),
            torch.as_tensor(M.data, dtype=torch.float64, device=DEV))


@pytest.mark.parametrize("values", [(1.0,), (1.0, 0.5), (1.0, 2 / 3, 1 / 3)])
def test_rap_terms_random_transfers(values):
    """C-ABI term list on random transfers P (1-3 entries a row, values from a
    set of 1-3) and a random symmetric A, R = P^T, against SciPy; coarse rows
    with no entries get the unit diagonal; 4 distinct values are refused."""
    import ctypes
    import scipy.sparse as sp
    from fealpy_amd import _native as nat
    rng = np.random.default_rng(len(values))
    nf, nc = 700, 150
    rows, cols, vals = [], [], []
    for i in range(nf):
        for J in rng.choice(nc - 10, size=rng.integers(1, 4), replace=False):  # last 10 coarse rows empty
            rows.append(i), cols.append(J), vals.append(values[rng.integers(0, len(values))])
    P = sp.csr_matrix((vals, (rows, cols)), shape=(nf, nc))
    B = sp.random(nf, nf, density=8 / nf, random_state=3, format="csr")
    A = (B + B.T + sp.eye(nf)).tocsr()
    R = P.T.tocsr()
    ref = (R @ A @ P).tocsr()
    ref = ref + sp.diags((np.diff(ref.indptr) == 0).astype(float))
    ref.sort_indices()
    L = nat.lib()
    r, a, p = _csr_dev(R), _csr_dev(A), _csr_dev(P)
    cand = int(np.diff(R.indptr).max() * np.diff(A.indptr).max() * np.diff(P.indptr).max())
    cptr = torch.empty(nc + 1, dtype=torch.int32, device=DEV)
    ws = nat.workspace(max(L.fa_rap_work_bytes(nc), L.fa_rap_terms_work_bytes(ref.nnz)), DEV)
    nnz, mrow = ctypes.c_int64(0), ctypes.c_int32(0)
    nat.check(L.fa_rap_count(nc, nc, nat.ptr(r[0]), nat.ptr(r[1]), nat.ptr(a[0]), nat.ptr(a[1]), nat.ptr(p[0]),
                             nat.ptr(p[1]), cand, nat.ptr(cptr), ctypes.byref(nnz), ctypes.byref(mrow), nat.ptr(ws),
                             ws.numel(), nat.stream()), "fa_rap_count")
    assert nnz.value == ref.nnz
    cidx = torch.empty(ref.nnz, dtype=torch.int32, device=DEV)
    nat.check(L.fa_rap_fill(nc, nat.ptr(r[0]), nat.ptr(r[1]), nat.ptr(a[0]), nat.ptr(a[1]), nat.ptr(p[0]),
                            nat.ptr(p[1]), cand, nat.ptr(cptr), nat.ptr(cidx), nat.stream()), "fa_rap_fill")
    wset = (ctypes.c_double * len(values))(*values)
    args = (nc, nat.ptr(r[0]), nat.ptr(r[1]), nat.ptr(r[2]), nat.ptr(a[0]), nat.ptr(a[1]), nat.ptr(p[0]),
            nat.ptr(p[1]), nat.ptr(p[2]), cand, nat.ptr(cptr), ref.nnz, wset, len(values))
    tptr = torch.empty(ref.nnz + 1, dtype=torch.int32, device=DEV)
    nt = ctypes.c_int64(0)
    nat.check(L.fa_rap_terms_count(*args, nat.ptr(tptr), ctypes.byref(nt), nat.ptr(ws), ws.numel(),
                                   nat.stream()), "fa_rap_terms_count")
    terms = torch.empty(nt.value, dtype=torch.int32, device=DEV)
    nat.check(L.fa_rap_terms_fill(*args, nat.ptr(tptr), nt.value, nat.ptr(terms), nat.ptr(ws), ws.numel(),
                                  nat.stream()), "fa_rap_terms_fill")
    cval = torch.empty(ref.nnz, dtype=torch.float64, device=DEV)
    nat.check(L.fa_rap_values_terms(nat.FA_F64, ref.nnz, nat.ptr(tptr), nat.ptr(terms), wset, len(values), nat.ptr(a[2]),
                                    nat.ptr(cval), nat.stream()), "fa_rap_values_terms")
    assert np.array_equal(cptr.cpu().numpy(), ref.indptr) and np.array_equal(cidx.cpu().numpy(), ref.indices)
    np.testing.assert_allclose(cval.cpu().numpy(), ref.data, rtol=0, atol=1e-13 * abs(ref.data).max())
    # entry offsets: contiguous lists, at least one term an entry (the unit
    # diagonal of an empty row), the last offset the term count
    tp = tptr.cpu().numpy()
    assert tp[0] == 0 and tp[-1] == nt.value and (np.diff(tp) >= 1).all()
    again = torch.empty_like(cval)
    nat.check(L.fa_rap_values_terms(nat.FA_F64, ref.nnz, nat.ptr(tptr), nat.ptr(terms), wset, len(values),
                                    nat.ptr(a[2]), nat.ptr(again), nat.stream()), "fa_rap_values_terms")
    assert torch.equal(again, cval)                                   # run to run: bitwise
    # coarse rows of 64 or more entries: past the staged form's bounds
    terms16 = torch.empty(nt.value, dtype=torch.int16, device=DEV)
    assert np.diff(ref.indptr).max() >= 64
    st = L.fa_rap_terms_fill_staged(*args, nat.ptr(tptr), nt.value, nat.ptr(terms16), nat.ptr(ws), ws.numel(),
                                    nat.stream())
    assert st == nat.FA_ECAPACITY
    four = (ctypes.c_double * 4)(1.0, 0.75, 0.5, 0.25)
    st = L.fa_rap_values_terms(nat.FA_F64, ref.nnz, nat.ptr(tptr), nat.ptr(terms), four, 4, nat.ptr(a[2]), nat.ptr(cval),
                               nat.stream())
    assert st == nat.FA_ECAPACITY


@pytest.mark.parametrize("values", [(1.0,), (1.0, 0.5), (1.0, 2 / 3, 1 / 3)])
def test_rap_staged_random_banded(values):
    """Staged words (fa_rap_terms_fill_staged / fa_rap_values_staged) on a
    random banded A and local random transfers, against SciPy and bit for bit
    against the 32-bit words; FP64 and FP32 operators."""
    import ctypes
    import scipy.sparse as sp
    from fealpy_amd import _native as nat
    rng = np.random.default_rng(10 + len(values))
    nf, nc = 600, 120
    rows, cols, vals = [], [], []
    for i in range(nf):
        near = [J for J in (i // 5 - 1, i // 5, i // 5 + 1) if 0 <= J < nc - 4]    # last coarse rows empty
        for J in rng.choice(near, size=min(len(near), int(rng.integers(1, 4))), replace=False):
            rows.append(i), cols.append(J), vals.append(values[rng.integers(0, len(values))])
    P = sp.csr_matrix((vals, (rows, cols)), shape=(nf, nc))
    B = sp.diags([rng.standard_normal(nf - k) for k in range(5)], list(range(5)), shape=(nf, nf))
    A = (B + B.T).tocsr()
    R = P.T.tocsr()
    ref = (R @ A @ P).tocsr()
    ref = ref + sp.diags((np.diff(ref.indptr) == 0).astype(float))
    ref.sort_indices()
    L = nat.lib()
    r, a, p = _csr_dev(R), _csr_dev(A), _csr_dev(P)
    cand = int(np.diff(R.indptr).max() * np.diff(A.indptr).max() * np.diff(P.indptr).max())
    cptr = torch.empty(nc + 1, dtype=torch.int32, device=DEV)
    ws = nat.workspace(max(L.fa_rap_work_bytes(nc), L.fa_rap_terms_work_bytes(ref.nnz)), DEV)
    nnz, mrow = ctypes.c_int64(0), ctypes.c_int32(0)
    nat.check(L.fa_rap_count(nc, nc, nat.ptr(r[0]), nat.ptr(r[1]), nat.ptr(a[0]), nat.ptr(a[1]), nat.ptr(p[0]),
                             nat.ptr(p[1]), cand, nat.ptr(cptr), ctypes.byref(nnz), ctypes.byref(mrow), nat.ptr(ws),
                             ws.numel(), nat.stream()), "fa_rap_count")
    assert nnz.value == ref.nnz and np.diff(ref.indptr).max() < 64
    cidx = torch.empty(ref.nnz, dtype=torch.int32, device=DEV)
    nat.check(L.fa_rap_fill(nc, nat.ptr(r[0]), nat.ptr(r[1]), nat.ptr(a[0]), nat.ptr(a[1]), nat.ptr(p[0]),
                            nat.ptr(p[1]), cand, nat.ptr(cptr), nat.ptr(cidx), nat.stream()), "fa_rap_fill")
    wset = (ctypes.c_double * len(values))(*values)
    args = (nc, nat.ptr(r[0]), nat.ptr(r[1]), nat.ptr(r[2]), nat.ptr(a[0]), nat.ptr(a[1]), nat.ptr(p[0]),
            nat.ptr(p[1]), nat.ptr(p[2]), cand, nat.ptr(cptr), ref.nnz, wset, len(values))
    tptr = torch.empty(ref.nnz + 1, dtype=torch.int32, device=DEV)
    nt = ctypes.c_int64(0)
    nat.check(L.fa_rap_terms_count(*args, nat.ptr(tptr), ctypes.byref(nt), nat.ptr(ws), ws.numel(),
                                   nat.stream()), "fa_rap_terms_count")
    terms = torch.empty(nt.value, dtype=torch.int32, device=DEV)
    nat.check(L.fa_rap_terms_fill(*args, nat.ptr(tptr), nt.value, nat.ptr(terms), nat.ptr(ws), ws.numel(),
                                  nat.stream()), "fa_rap_terms_fill")
    terms16 = torch.empty(nt.value, dtype=torch.int16, device=DEV)
    nat.check(L.fa_rap_terms_fill_staged(*args, nat.ptr(tptr), nt.value, nat.ptr(terms16), nat.ptr(ws), ws.numel(),
                                         nat.stream()), "fa_rap_terms_fill_staged")
    for dt, code in ((torch.float64, nat.FA_F64), (torch.float32, nat.FA_F32)):
        av = a[2].to(dt)
        glob = torch.empty(ref.nnz, dtype=dt, device=DEV)
        nat.check(L.fa_rap_values_terms(code, ref.nnz, nat.ptr(tptr), nat.ptr(terms), wset, len(values), nat.ptr(av),
                                        nat.ptr(glob), nat.stream()), "fa_rap_values_terms")
        staged = torch.empty_like(glob)
        nat.check(L.fa_rap_values_staged(code, nc, nat.ptr(r[0]), nat.ptr(r[1]), nat.ptr(a[0]), nat.ptr(cptr),
                                         nat.ptr(tptr), nat.ptr(terms16), wset, len(values), nat.ptr(av),
                                         nat.ptr(staged), nat.stream()), "fa_rap_values_staged")
        assert torch.equal(staged, glob), dt
    np.testing.assert_allclose(staged.double().cpu().numpy(), ref.data, rtol=0,
                               atol=1e-6 * abs(ref.data).max())


@pytest.mark.parametrize("p", [1, 2])
def test_rap_staged_equals_global_terms(p, monkeypatch):
    """Every level's Galerkin product through the staged words (LDS slots)
    equals the one through 32-bit words with A entry indices, bit for bit,
    in FP64 and FP32."""
    from fealpy_amd.solver import mg as mgmod
    n = {1: 10, 2: 6}[p]
    mesh, space, bc, A, b = _system(n, p=p)
    k = _lognormal(mesh, seed=5)
    _, _, _, A, _ = _system(n, k=k, p=p)
    for dt in (torch.float64, torch.float32):
        Ad = A if dt == torch.float64 else _as_type(A, dt)
        out = {}
        for staged in (True, False):
            monkeypatch.setattr(mgmod, "RAP_STAGED", staged)
            mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float64, coarse_max=1)
            mg.setup(Ad)
            kinds = [s[-1][-1] for lv in mg.levels[:-1] for s in lv.stages if s[-1] is not None]
            assert all(kk == staged for kk in kinds) and kinds
            out[staged] = [lv.A.values.clone() for lv in mg.levels[1:]]
        for l, (x1, x2) in enumerate(zip(out[True], out[False])):
            assert x1.dtype == dt and torch.equal(x1, x2), f"level {l + 1} {dt}"


def _as_type(A, dt):
    from fealpy_amd.sparse import CSRTensor
    return CSRTensor(A.crow, A.col, A.values.to(dt), A.shape)


def test_rap_staged_refuses_wide_rows():
    """A coarse row whose fine rows hold more than 512 entries is refused by
    the staged fill (FA_ECAPACITY); the 32-bit words take it."""
    import ctypes
    import scipy.sparse as sp
    from fealpy_amd import _native as nat
    nf, nc = 200, 4
    P = sp.csr_matrix((np.ones(nf), (np.arange(nf), np.arange(nf) % nc)), shape=(nf, nc))
    A = (sp.random(nf, nf, density=0.2, random_state=1, format="csr") + sp.eye(nf)).tocsr()
    R = P.T.tocsr()
    L = nat.lib()
    r, a, p = _csr_dev(R), _csr_dev(A), _csr_dev(P)
    cand = int(np.diff(R.indptr).max() * np.diff(A.indptr).max() * np.diff(P.indptr).max())
    assert cand <= 8192
    cptr = torch.empty(nc + 1, dtype=torch.int32, device=DEV)
    ref = (R @ A @ P).tocsr()
    ws = nat.workspace(max(L.fa_rap_work_bytes(nc), L.fa_rap_terms_work_bytes(ref.nnz)), DEV)
    nnz, mrow = ctypes.c_int64(0), ctypes.c_int32(0)
    nat.check(L.fa_rap_count(nc, nc, nat.ptr(r[0]), nat.ptr(r[1]), nat.ptr(a[0]), nat.ptr(a[1]), nat.ptr(p[0]),
                             nat.ptr(p[1]), cand, nat.ptr(cptr), ctypes.byref(nnz), ctypes.byref(mrow), nat.ptr(ws),
                             ws.numel(), nat.stream()), "fa_rap_count")
    cidx = torch.empty(nnz.value, dtype=torch.int32, device=DEV)
    nat.check(L.fa_rap_fill(nc, nat.ptr(r[0]), nat.ptr(r[1]), nat.ptr(a[0]), nat.ptr(a[1]), nat.ptr(p[0]),
                            nat.ptr(p[1]), cand, nat.ptr(cptr), nat.ptr(cidx), nat.stream()), "fa_rap_fill")
    wset = (ctypes.c_double * 1)(1.0)
    args = (nc, nat.ptr(r[0]), nat.ptr(r[1]), nat.ptr(r[2]), nat.ptr(a[0]), nat.ptr(a[1]), nat.ptr(p[0]),
            nat.ptr(p[1]), nat.ptr(p[2]), cand, nat.ptr(cptr), nnz.value, wset, 1)
    tptr = torch.empty(nnz.value + 1, dtype=torch.int32, device=DEV)
    nt = ctypes.c_int64(0)
    nat.check(L.fa_rap_terms_count(*args, nat.ptr(tptr), ctypes.byref(nt), nat.ptr(ws), ws.numel(), nat.stream()),
              "fa_rap_terms_count")
    terms16 = torch.empty(nt.value, dtype=torch.int16, device=DEV)
    st = L.fa_rap_terms_fill_staged(*args, nat.ptr(tptr), nt.value, nat.ptr(terms16), nat.ptr(ws), ws.numel(),
                                    nat.stream())
    assert st == nat.FA_ECAPACITY and b"staged" in L.fa_last_error()
    terms = torch.empty(nt.value, dtype=torch.int32, device=DEV)
    nat.check(L.fa_rap_terms_fill(*args, nat.ptr(tptr), nt.value, nat.ptr(terms), nat.ptr(ws), ws.numel(),
                                  nat.stream()), "fa_rap_terms_fill")
    cval = torch.empty(nnz.value, dtype=torch.float64, device=DEV)
    nat.check(L.fa_rap_values_terms(nat.FA_F64, nnz.value, nat.ptr(tptr), nat.ptr(terms), wset, 1, nat.ptr(a[2]),
                                    nat.ptr(cval), nat.stream()), "fa_rap_values_terms")
    ref.sort_indices()
    np.testing.assert_allclose(cval.cpu().numpy(), ref.data, rtol=1e-13)


def test_rap_term_list_elasticity():
    """Vector P2 -> P1 transfer (per component) through the term list."""
    import scipy.sparse as sp
    n = 4
    mesh = fa.TetrahedronMesh.from_box([0, 1, 0, 1, 0, 1], nx=n, ny=n, nz=n, device=DEV)
    ts = fa.TensorFunctionSpace(fa.LagrangeFESpace(mesh, 2), (-1, 3))
    mat = fa.LinearElasticMaterial("E", lame_lambda=1.0, shear_modulus=0.5)
    A = fa.BilinearForm(ts).add_integrator(fa.LinearElasticityIntegrator(mat, q=3)).assembly()
    bc = fa.DirichletBC(ts, gd=0.0)
    A, _ = bc.apply(A, torch.zeros(A.shape[0], dtype=torch.float64, device=DEV))
    mg = GeometricMG(ts, bc.is_boundary_dof, dtype=torch.float64, coarse_max=1)
    mg.setup(A)
    f, c = mg.levels[0], mg.levels[1]
    assert len(f.stages) == 1 and f.stages[0][-1] is not None
    ref = (f.R.to_scipy() @ f.A.to_scipy() @ f.P.to_scipy()).tocsr()
    ref = ref + sp.diags((np.diff(ref.indptr) == 0).astype(float))
    assert abs(c.A.to_scipy() - ref).max() < 1e-12 * abs(ref).max()


@pytest.mark.parametrize("p,n", [(3, 4), (4, 3)])
def test_mg_pcg_p3_p4_two_products(p, n):
    """P3/P4 -> P1 in 3D: the product of the row maxima bounds R A P at
    34,125 / 258,300 candidates per row; the hierarchy counts the exact
    candidates and forms R A P at once when they fit the symbolic kernel,
    else A P, then R (A P).  Every level equals SciPy's triple product (unit
    diagonal on empty rows); the MG-preconditioned CG reaches the tight
    solution."""
    import scipy.sparse as sp
    mesh = fa.TetrahedronMesh.from_box(nx=n, ny=n, nz=n, device=DEV)
    space = fa.LagrangeFESpace(mesh, p)
    k = _lognormal(mesh, seed=13)
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=p + 2)).assembly()
    b = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(1.0, q=p + 2)).assembly()
    bc = fa.DirichletBC(space, gd=0.0)
    A, b = bc.apply(A, b)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float64)
    mg.setup(A)
    from fealpy_amd.solver.mg import RAP_CANDIDATES, _candidates
    f0 = mg.levels[0]
    assert len(f0.stages) == (1 if _candidates(f0.R, f0.A, f0.P) <= RAP_CANDIDATES else 2)
    for l in range(len(mg.levels) - 1):
        f, c = mg.levels[l], mg.levels[l + 1]
        ref = (f.R.to_scipy() @ f.A.to_scipy() @ f.P.to_scipy()).tocsr()
        ref = ref + sp.diags((np.diff(ref.indptr) == 0).astype(float))
        got = c.A.to_scipy()
        assert abs(got - ref).max() < 1e-12 * abs(ref).max(), l
    x_ref = fa.cg(A, b, atol=1e-30, rtol=1e-13, maxit=50000)
    for dt in (torch.float64, torch.float32):
        m = GeometricMG(space, bc.is_boundary_dof, dtype=dt)
        x, info = fa.cg(A, b, M=m, atol=1e-30, rtol=1e-11, maxit=300, returninfo=True)
        err = float(torch.linalg.norm(x - x_ref)) / float(torch.linalg.norm(x_ref))
        assert err < 1e-8, (dt, err, info)
        assert info["niter"] < 100, info


def test_sell_padding_columns_in_range():
    """Padding entries of rectangular transfers (P: fine x coarse) must point at
    a valid column of the operand (regression: they used the row index)."""
    mesh, space, bc, A, b = _system(16)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float64)
    for lv in mg.levels[:-1]:
        for S in (lv.Ps, lv.Rs):
            assert int(S.cols[:max(S.entries, 1)].max()) < S.shape[1]
            assert int(S.cols[:max(S.entries, 1)].min()) >= 0


def test_vcycle_is_symmetric_and_spd():
    mesh, space, bc, A, b = _system(16)
    k = _lognormal(mesh)
    _, _, _, A, _ = _system(16, k=k)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float64).setup(A)
    g = torch.Generator(device="cpu").manual_seed(0)
    r1 = torch.randn(A.shape[0], generator=g, dtype=torch.float64).to(DEV)
    r2 = torch.randn(A.shape[0], generator=g, dtype=torch.float64).to(DEV)
    z1, z2 = mg @ r1, mg @ r2
    a, bb = float(torch.dot(z1, r2)), float(torch.dot(r1, z2))
    assert abs(a - bb) < 1e-10 * max(abs(a), abs(bb))
    assert float(torch.dot(z1, r1)) > 0


@pytest.mark.parametrize("n", [16, 32])
def test_mg_pcg_fp64_and_mixed(n):
    mesh, space, bc, A, b = _system(n)
    k = _lognormal(mesh)
    _, _, _, A, b = _system(n, k=k)
    x_ref, info_ref = fa.cg(A, b, atol=1e-30, rtol=1e-13, maxit=50000, returninfo=True)
    nrm = float(torch.linalg.norm(x_ref))
    for dtype in (torch.float64, torch.float32):
        mg = GeometricMG(space, bc.is_boundary_dof, dtype=dtype)
        x, info = fa.cg(A, b, M=mg, atol=1e-30, rtol=1e-11, maxit=200, returninfo=True)
        err = float(torch.linalg.norm(x - x_ref)) / nrm
        assert err < 1e-8, (dtype, err, info)
        assert info["niter"] < 60, (dtype, info, info_ref)
        # FP32 cycles smooth the finest level with FP16 D^-1 A and int16 column offsets
        assert getattr(mg.levels[0], "compressed", False) == (dtype == torch.float32)


def test_compressed_level_matches_full_format():
    """The compressed smoother operator (FP16 D^-1 A, int16 offsets) applied
    to a vector equals D^-1 A x within FP16 rounding, and a V-cycle with it
    is symmetric to that rounding (a_ij / a_ii and a_ji / a_jj round
    separately, so exact symmetry is not kept)."""
    from fealpy_amd.sparse.sell import SellMatrix
    mesh, space, bc, A, b = _system(12)
    k = _lognormal(mesh, seed=9)
    _, _, _, A, b = _system(12, k=k)
    S = SellMatrix(A.crow, A.col, A.shape)
    ws = torch.empty(64, dtype=torch.uint8, device=DEV)
    assert S.set_compressed(A.values, ws)
    offs = S.h_offs.long()
    rows = torch.arange(S.entries, device=DEV)
    sl = torch.searchsorted(S.slice_ptr[1:].long(), rows, right=True)
    row = sl * 64 + (rows - S.slice_ptr[sl].long()) % 64
    x = torch.rand(A.shape[0], dtype=torch.float64, device=DEV)
    valid = row < A.shape[0]
    y = torch.zeros(A.shape[0], dtype=torch.float64, device=DEV)
    y.index_add_(0, row[valid], S.h_vals.double()[valid] * x[(row + offs)[valid]])
    d = A.diagonal()
    ref = (A @ x) / d
    assert float((y - ref).abs().max()) < 2e-3 * float(ref.abs().max())
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32).setup(A)
    g = torch.Generator(device="cpu").manual_seed(1)
    r1 = torch.randn(A.shape[0], generator=g).to(DEV)
    r2 = torch.randn(A.shape[0], generator=g).to(DEV)
    z1, z2 = mg @ r1, mg @ r2
    a, bb = float(torch.dot(z1.double(), r2.double())), float(torch.dot(r1.double(), z2.double()))
    assert abs(a - bb) < 1e-3 * max(abs(a), abs(bb))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_smoother_setup_matches_csr_setup(dtype):
    """fa_smoother_setup (one pass over the FP64 SELL) gives the dinv of
    fa_jacobi_setup bit for bit, its Gershgorin bound to rounding, 1/dinv as
    diag, the FP16 D^-1 A of fa_sell16_fill and the FP32 cast of the values."""
    import ctypes
    from fealpy_amd import _native as nat
    from fealpy_amd.sparse.sell import FaSell16
    mesh, space, bc, A, b = _system(10)
    k = _lognormal(mesh, seed=4)
    _, _, _, A, b = _system(10, k=k)
    L = nat.lib()
    n = A.shape[0]
    code = nat.dtype_code(dtype)
    ws = nat.workspace(L.fa_reduce_work_bytes(), DEV)
    dinv_ref = torch.empty(n, dtype=dtype, device=DEV)
    lam = ctypes.c_double(0.0)
    nat.check(L.fa_jacobi_setup(code, n, A.nnz, nat.ptr(A.crow), nat.ptr(A.col), nat.ptr(A.values),
                                nat.ptr(dinv_ref), ctypes.byref(lam), nat.ptr(ws), ws.numel(), nat.stream()), "jacobi")
    S = SellMatrix(A.crow, A.col, A.shape)
    S.set_values(A.values, torch.float64)
    wsc = torch.empty(64, dtype=torch.uint8, device=DEV)
    assert S.set_compressed(A.values, wsc)
    h_ref = S.h_vals.clone()
    S.h_vals.zero_()
    v32 = torch.full((S.entries,), 7.0, dtype=torch.float32, device=DEV)
    dinv = torch.empty(n, dtype=dtype, device=DEV)
    diag = torch.empty(n, dtype=dtype, device=DEV)
    bound = torch.zeros(1, dtype=torch.float64, device=DEV)
    d64, h16 = S.desc(torch.float64), S.desc16()
    nat.check(L.fa_smoother_setup(code, nat.FA_F64, ctypes.byref(d64), nat.ptr(dinv), nat.ptr(diag), nat.ptr(v32),
                                  ctypes.byref(h16), nat.ptr(bound), nat.ptr(ws), ws.numel(), nat.stream()),
              "fa_smoother_setup")
    assert torch.equal(dinv, dinv_ref)
    assert torch.equal(diag, torch.reciprocal(dinv))
    assert abs(float(bound[0]) - lam.value) <= 1e-14 * lam.value
    assert torch.equal(S.h_vals, h_ref)
    assert torch.equal(v32, S.vals[torch.float64].float())


def test_mg_pcg_2d():
    mesh, space, bc, A, b = _system(64, dim=2)
    x_ref = fa.cg(A, b, atol=1e-30, rtol=1e-13, maxit=50000)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32)
    x, info = fa.cg(A, b, M=mg, atol=1e-30, rtol=1e-11, maxit=200, returninfo=True)
    assert float(torch.linalg.norm(x - x_ref)) / float(torch.linalg.norm(x_ref)) < 1e-8
    assert info["niter"] < 40


@pytest.mark.parametrize("n,dim,p", [(45, 2, 1), (45, 2, 2), (27, 3, 1), (24, 3, 1)])
def test_mg_pcg_odd_counts(n, dim, p):
    """Odd cell counts coarsen to (n+1)/2 on non-nested grids (transfer =
    coarse P1 function at the fine nodes) down to a dense coarsest level; the
    iteration count stays that of nested hierarchies."""
    mesh, space, bc, A, b = _system(n, dim=dim, p=p)
    x_ref = fa.cg(A, b, atol=1e-30, rtol=1e-13, maxit=50000)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32)
    assert mg.coarse_dense and mg.levels[-1].n <= 64
    x, info = fa.cg(A, b, M=mg, atol=1e-30, rtol=1e-11, maxit=200, returninfo=True)
    assert float(torch.linalg.norm(x - x_ref)) / float(torch.linalg.norm(x_ref)) < 1e-8
    assert info["niter"] < 30, info


def test_mg_polynomial_coarsest_level():
    """A coarsest level above 2048 rows (coarse_max above that) is smoothed
    from zero by a Chebyshev polynomial instead of inverted."""
    mesh, space, bc, A, b = _system(63, dim=2)
    x_ref = fa.cg(A, b, atol=1e-30, rtol=1e-13, maxit=50000)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32, coarse_max=4096)
    assert not mg.coarse_dense and len(mg.levels) == 1
    x, info = fa.cg(A, b, M=mg, atol=1e-30, rtol=1e-11, maxit=400, returninfo=True)
    assert float(torch.linalg.norm(x - x_ref)) / float(torch.linalg.norm(x_ref)) < 1e-8


@pytest.mark.parametrize("dim,dt", [(2, torch.float32), (3, torch.float32), (3, torch.float64)])
def test_pcg_graph_replays_equal_eager(dim, dt):
    """fa_pcg_graph (iterations replayed from a HIP graph) gives bitwise the
    eager fa_pcg result: over several solves in a row, with eager work and
    synchronisations between them, after a re-setup with new values (new
    smoother bounds: the executable graph is updated), and with maxit not a
    multiple of check_every."""
    n = 48 if dim == 2 else 12
    mesh, space, bc, A, b = _system(n, dim=dim)
    mg_e = GeometricMG(space, bc.is_boundary_dof, dtype=dt)
    mg_g = GeometricMG(space, bc.is_boundary_dof, dtype=dt, graphs=True)
    x_e, info_e = fa.cg(A, b, M=mg_e, atol=1e-30, rtol=1e-10, maxit=200, returninfo=True)
    for rep in range(3):
        x_g, info_g = fa.cg(A, b, M=mg_g, atol=1e-30, rtol=1e-10, maxit=200, returninfo=True)
        assert torch.equal(x_g, x_e), rep
        assert info_g == info_e, (rep, info_g, info_e)
        y = fa.cg(A, 2 * b, M=mg_e, atol=1e-30, rtol=1e-6, maxit=200)      # eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(y).all()
    k = _lognormal(mesh, seed=3)
    _, _, _, A2, b2 = _system(n, k=k, dim=dim)
    x_e2, info_e2 = fa.cg(A2, b2, M=mg_e, atol=1e-30, rtol=1e-10, maxit=300, returninfo=True)
    x_g2, info_g2 = fa.cg(A2, b2, M=mg_g, atol=1e-30, rtol=1e-10, maxit=300, returninfo=True)
    assert torch.equal(x_g2, x_e2) and info_g2 == info_e2
    x_e3, info_e3 = fa.cg(A2, b2, M=mg_e, atol=1e-30, rtol=1e-14, maxit=7, returninfo=True)
    x_g3, info_g3 = fa.cg(A2, b2, M=mg_g, atol=1e-30, rtol=1e-14, maxit=7, returninfo=True)
    assert info_e3["niter"] == 7 and info_g3 == info_e3 and torch.equal(x_g3, x_e3)


def test_cg_batch_matches_oracle():
    """Right-hand sides as columns (fealpy/solver/cg.py:69-123): lockstep
    iterations, alpha and beta per column, one stopping test on the summed
    r.z -- the oracle's iteration count and iterates; batch_first gives the
    transposed result."""
    from oracle import fem as ofem
    mesh, space, bc, A, b = _system(12, dim=2)
    Bm = torch.as_tensor(np.random.default_rng(6).standard_normal((A.shape[0], 3)) * np.array([1.0, 10.0, 0.1]),
                         device=DEV)
    X, info = fa.cg(A, Bm, atol=0.0, rtol=1e-8, maxit=500, returninfo=True)
    Xo, info_o = ofem.cg(A.to_scipy().tocsr(), Bm.cpu().numpy(), atol=0.0, rtol=1e-8, maxit=500)
    assert info["niter"] == info_o["niter"]
    np.testing.assert_allclose(X.cpu().numpy(), Xo, rtol=0, atol=1e-9 * np.abs(Xo).max())
    Xt = fa.cg(A, Bm.t().contiguous(), batch_first=True, atol=0.0, rtol=1e-8, maxit=500)
    assert torch.equal(Xt, X.t())


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_cg_batch_with_multigrid_and_generic_operator(dt):
    """A batch through the multigrid preconditioner (either cycle precision)
    and a 1-D solve through a generic operator (anything with @) with the
    multigrid cycle: every column reaches its tight single solve."""
    mesh, space, bc, A, b = _system(16, dim=2)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=dt)
    Bm = torch.stack([b, 2 * b + 1e-3 * torch.arange(b.shape[0], device=DEV, dtype=torch.float64)], dim=1)
    Bm[bc.is_boundary_dof] = b[bc.is_boundary_dof, None]
    X, info = fa.cg(A, Bm, M=mg, atol=1e-30, rtol=1e-11, maxit=200, returninfo=True)
    assert info["niter"] < 40, info
    for k in range(2):
        ref = fa.cg(A, Bm[:, k].contiguous(), atol=1e-30, rtol=1e-13, maxit=50000)
        assert float(torch.linalg.norm(X[:, k] - ref)) < 1e-8 * float(torch.linalg.norm(ref)), k

    class Op:
        def __matmul__(self, v):
            return A @ v
    x, info = fa.cg(Op(), b, M=mg, atol=1e-30, rtol=1e-11, maxit=200, returninfo=True)
    ref = fa.cg(A, b, atol=1e-30, rtol=1e-13, maxit=50000)
    assert float(torch.linalg.norm(x - ref)) < 1e-8 * float(torch.linalg.norm(ref))
    assert info["niter"] < 40, info


def test_pcg_reproducible():
    mesh, space, bc, A, b = _system(16)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32)
    x1 = fa.cg(A, b, M=mg, rtol=1e-10, maxit=200)
    x2 = fa.cg(A, b, M=mg, rtol=1e-10, maxit=200)
    assert torch.equal(x1, x2)


def test_pcg_result_independent_of_host_checks():
    """The host looks at the PCG state after a predicted number of
    iterations (at most check_every): the iterate, the iteration count and
    the residual are the same whatever the interval, with and without the
    preconditioner, and maxit still stops at exactly maxit."""
    mesh, space, bc, A, b = _system(16)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32)
    runs = [mg.pcg(A, b, atol=0.0, rtol=1e-10, maxit=200, check_every=c) for c in (1, 3, 16, 64)]
    for x, info in runs[1:]:
        assert torch.equal(x, runs[0][0]) and info == runs[0][1]
    plain = [fa.cg(A, b, atol=0.0, rtol=1e-9, maxit=2000, returninfo=True, check_every=c) for c in (1, 7, 50)]
    for x, info in plain[1:]:
        assert torch.equal(x, plain[0][0]) and info == plain[0][1]
    x, info = mg.pcg(A, b, atol=0.0, rtol=1e-14, maxit=5, check_every=16)
    assert info["niter"] == 5


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_pcg_from_zero_equals_zero_start(dt):
    """fa_pcg_from_zero (MG.pcg with no x0) against fa_pcg started from an
    explicit zero vector: the same iterate, iteration count and residual,
    bit for bit, whatever x held before; a zero right-hand side gives x = 0."""
    mesh, space, bc, A, b = _system(12)
    from fealpy_amd.sparse import CSRTensor
    Ad = A if dt == torch.float64 else CSRTensor(A.crow, A.col, A.values.to(dt), A.shape)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32)
    x1, i1 = mg.pcg(Ad, b, atol=0.0, rtol=1e-9, maxit=200)
    x2, i2 = mg.pcg(Ad, b, x0=torch.zeros_like(b), atol=0.0, rtol=1e-9, maxit=200)
    assert torch.equal(x1, x2) and i1 == i2
    z, iz = mg.pcg(Ad, torch.zeros_like(b), atol=0.0, rtol=1e-9, maxit=200)
    assert float(z.abs().max()) == 0.0 and iz["niter"] == 0


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_grid_transfers_equal_sell_transfers(dt, monkeypatch):
    """Nested 3D levels apply R and P from the grid (k_box_restrict,
    k_box_prolong): the PCG iterate, iteration count and residual equal those
    with the SELL transfers bit for bit, with a Dirichlet boundary, a
    lognormal coefficient and a level of odd counts (16 x 8 x 12 -> 8 x 4 x 6
    -> 4 x 2 x 3, whose next transfer keeps its SELL matrix)."""
    from fealpy_amd.solver import mg as mgmod
    mesh = fa.TetrahedronMesh.from_box(nx=16, ny=8, nz=12, device=DEV)
    u = fa.cartesian(lambda x: torch.sin(PI * x[..., 0]) * torch.sin(PI * x[..., 1]) * x[..., 2])
    space = fa.LagrangeFESpace(mesh, 1)
    k = _lognormal(mesh, seed=9)
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=2)).assembly()
    b = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(1.0, q=2)).assembly()
    bc = fa.DirichletBC(space, gd=u)
    A, b = bc.apply(A, b)
    out = {}
    for grid in (True, False):
        monkeypatch.setattr(mgmod, "GRID_TRANSFERS", grid)
        mg = GeometricMG(space, bc.is_boundary_dof, dtype=dt, coarse_max=8)
        assert mg.levels[0].grid_transfer == (16, 8, 12) and mg.levels[1].grid_transfer == (8, 4, 6)
        assert mg.levels[2].grid_transfer is None
        out[grid] = mg.pcg(A, b, atol=0.0, rtol=1e-10, maxit=200)
    (x1, i1), (x2, i2) = out[True], out[False]
    assert torch.equal(x1, x2) and i1 == i2, (i1, i2)


def test_cg_maxit_and_trivial_rhs():
    mesh, space, bc, A, b = _system(8)
    x, info = fa.cg(A, b, maxit=3, returninfo=True)
    assert info["niter"] == 3
    z = fa.cg(A, torch.zeros_like(b))
    assert float(z.abs().max()) == 0.0


def test_mg_pcg_p2_scalar():
    """P2 -> P1 (same mesh) -> geometric levels."""
    mesh, space, bc, A, b = _system(12, p=2)
    k = _lognormal(mesh, seed=7)
    _, _, _, A, b = _system(12, k=k, p=2)
    x_ref = fa.cg(A, b, atol=1e-30, rtol=1e-13, maxit=50000)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32)
    assert mg.levels[1].n == mesh.number_of_nodes()
    x, info = fa.cg(A, b, M=mg, atol=1e-30, rtol=1e-11, maxit=300, returninfo=True)
    err = float(torch.linalg.norm(x - x_ref)) / float(torch.linalg.norm(x_ref))
    assert err < 1e-8, (err, info)
    assert info["niter"] < 80, info


@pytest.mark.parametrize("shape", [(-1, 3), (3, -1)])
def test_mg_pcg_elasticity_p2(shape):
    """Vector P2 elasticity, clamped: P2 -> P1 -> geometric, per component."""
    n = 8
    mesh = fa.TetrahedronMesh.from_box([0, 1, 0, 1, 0, 1], nx=n, ny=n, nz=n, device=DEV)
    ts = fa.TensorFunctionSpace(fa.LagrangeFESpace(mesh, 2), shape)
    mat = fa.LinearElasticMaterial("E", lame_lambda=1.0, shear_modulus=0.5)
    A = fa.BilinearForm(ts).add_integrator(fa.LinearElasticityIntegrator(mat, q=3)).assembly()
    F = fa.LinearForm(ts).add_integrator(fa.VectorSourceIntegrator(
        torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64, device=DEV), q=3)).assembly()
    bc = fa.DirichletBC(ts, gd=0.0)
    A, F = bc.apply(A, F)
    x_ref = fa.cg(A, F, atol=1e-30, rtol=1e-13, maxit=50000)
    mg = GeometricMG(ts, bc.is_boundary_dof, dtype=torch.float32)
    assert mg.levels[1].n == 3 * mesh.number_of_nodes()
    x, info = fa.cg(A, F, M=mg, atol=1e-30, rtol=1e-11, maxit=300, returninfo=True)
    err = float(torch.linalg.norm(x - x_ref)) / float(torch.linalg.norm(x_ref))
    assert err < 1e-8, (err, info)
    assert info["niter"] < 120, info


@pytest.mark.parametrize("n", [40, 89, 90, 300, 648])
def test_dense_spd_inverse(n):
    """fa_dense_spd_inverse (LDS Gauss-Jordan up to 89 rows, blocked above)
    against numpy.linalg.inv of the same SPD matrix."""
    from fealpy_amd import _native as nat
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n))
    A = X @ X.T + n * np.eye(n)
    a = torch.as_tensor(A, device=DEV).contiguous()
    nat.check(nat.lib().fa_dense_spd_inverse(n, nat.ptr(a), nat.stream()), "fa_dense_spd_inverse")
    ref = np.linalg.inv(A)
    assert np.abs(a.cpu().numpy() - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_reused_hierarchy_equals_fresh(dtype):
    """A hierarchy set up again for a new coefficient equals a fresh one bit
    for bit: coarse operators, their SELL copies, smoother data and the
    V-cycle (regression: coarse SELL copies were not refilled after the
    Galerkin kernel rewrote the coarse values in place)."""
    mesh, space, bc, A, b = _system(16)
    k1, k2 = _lognormal(mesh, seed=21), _lognormal(mesh, seed=22)
    _, _, _, A1, _ = _system(16, k=k1)
    _, _, _, A2, b2 = _system(16, k=k2)
    used = GeometricMG(space, bc.is_boundary_dof, dtype=dtype).setup(A1)
    used.setup(A2)
    fresh = GeometricMG(space, bc.is_boundary_dof, dtype=dtype).setup(A2)
    for lu, lf in zip(used.levels, fresh.levels):
        assert torch.equal(lu.A.values, lf.A.values)
        assert torch.equal(lu.As.vals[torch.float64], lf.As.vals[torch.float64])
        assert torch.equal(lu.dinv, lf.dinv) and lu.lmax == lf.lmax
        if getattr(lf, "compressed", False):
            assert torch.equal(lu.As.h_vals, lf.As.h_vals)
    r = b2.to(dtype)
    assert torch.equal(used @ r, fresh @ r)
    x1, i1 = fa.cg(A2, b2, M=used, atol=0.0, rtol=1e-10, returninfo=True)
    x2, i2 = fa.cg(A2, b2, M=fresh, atol=0.0, rtol=1e-10, returninfo=True)
    assert torch.equal(x1, x2) and i1["niter"] == i2["niter"]


@pytest.mark.parametrize("p", [1, 3])
def test_reused_hierarchy_same_pattern_equals_fresh(p):
    """A new matrix on the SAME pattern (same space, new coefficient: the
    symbolic products are kept) gives the fresh hierarchy's coarse operators
    and solution bit for bit -- the Galerkin products read the new fine
    values, with one product (p = 1) and with two (P3 -> P1)."""
    n = 10 if p == 1 else 4
    mesh = fa.TetrahedronMesh.from_box(nx=n, ny=n, nz=n, device=DEV)
    space = fa.LagrangeFESpace(mesh, p)
    bc = fa.DirichletBC(space, gd=0.0)

    def system(seed):
        k = _lognormal(mesh, seed=seed)
        A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=p + 2)).assembly()
        b = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(1.0, q=p + 2)).assembly()
        return bc.apply(A, b)
    A1, _ = system(31)
    A2, b2 = system(32)
    assert A1.crow.data_ptr() == A2.crow.data_ptr()          # one cached pattern
    used = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32).setup(A1)
    used.setup(A2)
    fresh = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32).setup(A2)
    for lu, lf in zip(used.levels[1:], fresh.levels[1:]):
        assert torch.equal(lu.A.values, lf.A.values)
    x1, i1 = fa.cg(A2, b2, M=used, atol=0.0, rtol=1e-10, returninfo=True)
    x2, i2 = fa.cg(A2, b2, M=fresh, atol=0.0, rtol=1e-10, returninfo=True)
    assert torch.equal(x1, x2) and i1["niter"] == i2["niter"]


def test_fp16_operator_on_int32_columns_equals_offsets():
    """The FP16 smoother operator read through the level's int32 columns (the
    form used when a column lies beyond the int16 offset range, e.g. P2 DOF
    numbering on large meshes) gives the V-cycle and the MG-PCG solution of the
    int16-offset form bit for bit."""
    mesh, space, bc, A, b = _system(12)
    k = _lognormal(mesh, seed=17)
    _, _, _, A, b = _system(12, k=k)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32).setup(A)
    assert mg.levels[0].compressed and mg.levels[0].As.h_offs is not None
    r = b.float()
    z_off = mg @ r
    x_off, i_off = fa.cg(A, b, M=mg, atol=0.0, rtol=1e-10, returninfo=True)
    for lv in mg.levels:
        if getattr(lv, "compressed", False):
            lv.As.h_offs = None
            lv.As._h_band_ok = False
    mg.setup(A)
    assert mg.levels[0].compressed and mg._desc.levels[0].Ah.offs is None
    assert torch.equal(mg @ r, z_off)
    x_col, i_col = fa.cg(A, b, M=mg, atol=0.0, rtol=1e-10, returninfo=True)
    assert torch.equal(x_col, x_off) and i_col["niter"] == i_off["niter"]
