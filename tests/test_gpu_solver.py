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


def test_mg_pcg_2d():
    mesh, space, bc, A, b = _system(64, dim=2)
    x_ref = fa.cg(A, b, atol=1e-30, rtol=1e-13, maxit=50000)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32)
    x, info = fa.cg(A, b, M=mg, atol=1e-30, rtol=1e-11, maxit=200, returninfo=True)
    assert float(torch.linalg.norm(x - x_ref)) / float(torch.linalg.norm(x_ref)) < 1e-8
    assert info["niter"] < 40


def test_pcg_reproducible():
    mesh, space, bc, A, b = _system(16)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32)
    x1 = fa.cg(A, b, M=mg, rtol=1e-10, maxit=200)
    x2 = fa.cg(A, b, M=mg, rtol=1e-10, maxit=200)
    assert torch.equal(x1, x2)


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
