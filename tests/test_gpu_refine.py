"""Matrix-free P1 operator (BilinearForm.mult / residual), multigrid set up
from an FP32 operator, and the mixed-precision refinement solve (cg_refine).

Pinned against values written by running the reference: the lognormal P1
fixtures (3D, 8^3 and 24^3) hold its assembled matrix, its Dirichlet-applied
system and its tight CG solution."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import fealpy_amd as fa
from fealpy_amd.expr import cell, sin, x, y, z
from fealpy_amd.mesh import TetrahedronMesh, TriangleMesh
from fealpy_amd.solver import GeometricMG
from fealpy_amd.solver.refine import cg_refine

pytestmark = pytest.mark.gpu
PI = np.pi
DEV = torch.device("cuda")
U3 = sin(PI * x) * sin(PI * y) * sin(PI * z)


def _np(t):
    return t.detach().cpu().numpy()


def _problem(g):
    n = int(g["n"])
    k = torch.as_tensor(g["k"], device=DEV)
    mesh = TetrahedronMesh.from_box(nx=n, ny=n, nz=n, device=DEV)
    space = fa.LagrangeFESpace(mesh, 1)
    form = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=2))
    f = 3 * PI ** 2 * cell(k) * U3
    F = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(f, q=2)).assembly()
    return mesh, space, form, F


@pytest.mark.parametrize("name", ["fixtures_diffusion3d_p1_lognormal.npz", "fixtures_diffusion3d_p1_lognormal_24.npz"])
def test_matrix_free_equals_reference_matrix(golden, name):
    """mult(x) = A x with A the reference's assembled matrix (1e-12 of |A||x|)."""
    g = golden(name)
    _, space, form, _ = _problem(g)
    n = int(g["gdof"])
    A = sp.csr_matrix((g["A_data"], g["A_indices"], g["A_indptr"]), shape=(n, n))
    rng = np.random.default_rng(7)
    for _ in range(2):
        xv = rng.standard_normal(n)
        y = _np(form.mult(torch.as_tensor(xv, device=DEV)))
        ref = A @ xv
        scale = abs(A).max() * np.abs(xv).max()
        np.testing.assert_allclose(y, ref, rtol=0, atol=1e-12 * scale * 16)
    assert form._M is None                        # nothing was assembled


def test_matrix_free_is_run_to_run_identical(golden):
    g = golden("fixtures_diffusion3d_p1_lognormal_24.npz")
    _, space, form, _ = _problem(g)
    xv = torch.as_tensor(np.random.default_rng(3).standard_normal(int(g["gdof"])), device=DEV)
    a, b = form.mult(xv), form.mult(xv)
    assert torch.equal(a, b)


def test_matrix_free_mass_and_2d():
    """Diffusion + mass group on triangles and tetrahedra against the assembled matrix."""
    for mesh in (TriangleMesh.from_box([0, 1, 0, 1], 20, 13, device=DEV),
                 TetrahedronMesh.from_box(nx=5, ny=4, nz=6, device=DEV)):
        space = fa.LagrangeFESpace(mesh, 1)
        nc = mesh.number_of_cells()
        k = torch.as_tensor(np.exp(np.random.default_rng(5).standard_normal(nc)), device=DEV)
        form = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=3),
                                                     fa.ScalarMassIntegrator(2.5, q=3))
        A = form.assembly()
        xv = torch.as_tensor(np.random.default_rng(1).standard_normal(space.number_of_global_dofs()), device=DEV)
        mf = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=3),
                                                   fa.ScalarMassIntegrator(2.5, q=3))
        y = _np(mf.mult(xv))
        ref = _np(A @ xv)
        np.testing.assert_allclose(y, ref, rtol=0, atol=1e-13 * np.abs(ref).max() * 10)


def test_residual_form(golden):
    """residual(b, x, bd) = b - A x off the boundary, 0 on it, with its squared
    norm, in FP64 and stored as FP32."""
    g = golden("fixtures_diffusion3d_p1_lognormal_24.npz")
    _, space, form, F = _problem(g)
    n = int(g["gdof"])
    A = sp.csr_matrix((g["A_data"], g["A_indices"], g["A_indptr"]), shape=(n, n))
    bd = torch.as_tensor(g["is_bd"].astype(np.uint8), device=DEV)
    xv = np.random.default_rng(11).standard_normal(n)
    ref = g["F"] - A @ xv
    ref[g["is_bd"]] = 0.0
    r, nr2 = form.residual(torch.as_tensor(g["F"], device=DEV), torch.as_tensor(xv, device=DEV), bd)
    np.testing.assert_allclose(_np(r), ref, rtol=0, atol=1e-12 * np.abs(ref).max() * 16)
    assert abs(float(nr2) - ref @ ref) <= 1e-12 * (ref @ ref)
    r32, n32 = form.residual(torch.as_tensor(g["F"], device=DEV), torch.as_tensor(xv, device=DEV), bd,
                             dtype=torch.float32)
    assert r32.dtype == torch.float32 and float(n32) == float(nr2)
    np.testing.assert_allclose(_np(r32), ref, rtol=0, atol=1e-6 * np.abs(ref).max())


def test_mg_setup_from_fp32_operator(golden):
    """A hierarchy set up from the FP32 Dirichlet-applied matrix: coarse
    operators within FP32 rounding of the FP64 hierarchy's."""
    g = golden("fixtures_diffusion3d_p1_lognormal_24.npz")
    _, space, form, F = _problem(g)
    bc = fa.DirichletBC(space, gd=U3)
    A64 = bc.apply_matrix(form.assembly())
    A32 = bc.apply_matrix(fa.BilinearForm(space).add_integrator(
        fa.ScalarDiffusionIntegrator(torch.as_tensor(g["k"], device=DEV), q=2)).assembly(dtype=torch.float32))
    assert A32.dtype == torch.float32
    m64 = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32).setup(A64)
    m32 = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32).setup(A32)
    assert len(m64.levels) == len(m32.levels) > 2
    for l64, l32 in zip(m64.levels[1:], m32.levels[1:]):
        assert l32.A.values.dtype == torch.float32
        a, b = _np(l64.A.values), _np(l32.A.values).astype(np.float64)
        np.testing.assert_allclose(b, a, rtol=0, atol=2e-6 * np.abs(a).max())
        assert abs(l64.lmax - l32.lmax) <= 1e-5 * l64.lmax


@pytest.mark.parametrize("name", ["fixtures_diffusion3d_p1_lognormal.npz", "fixtures_diffusion3d_p1_lognormal_24.npz"])
def test_cg_refine_reaches_reference_solution(golden, name):
    """FP32 assembly + FP32 MG-PCG corrections + FP64 matrix-free residuals:
    the solution within 1e-8 of the reference's tight FP64 solution."""
    g = golden(name)
    _, space, form, F = _problem(g)
    bc = fa.DirichletBC(space, gd=U3)
    A32 = bc.apply_matrix(form.assembly(dtype=torch.float32))
    form._M = None                                # the FP64 operator stays matrix-free
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32).setup(A32)
    xs, info = cg_refine(form, F, A32, mg, bc, rtol=1e-10, returninfo=True)
    ref = g["x"]
    err = np.linalg.norm(_np(xs) - ref) / np.linalg.norm(ref)
    assert err <= 1e-8, (err, info)
    assert info["refinements"] >= 2 and info["residual"] < 1e-10 * np.linalg.norm(g["Fbc"])
    bd = g["is_bd"]
    np.testing.assert_allclose(_np(xs)[bd], ref[bd], rtol=0, atol=1e-14)   # boundary values stay the BC's


def test_filtered_fp32_copy_of_fp64_matrix(golden):
    """apply_matrix(A64, dtype=float32): the reference's Dirichlet-applied
    pattern, values equal to the FP64 filter's rounded once."""
    g = golden("fixtures_diffusion3d_p1_lognormal_24.npz")
    _, space, form, F = _problem(g)
    bc = fa.DirichletBC(space, gd=U3)
    A64 = form.assembly()
    B64 = bc.apply_matrix(A64)
    B32 = bc.apply_matrix(A64, dtype=torch.float32)
    assert B32.dtype == torch.float32
    np.testing.assert_array_equal(_np(B32.col), g["Abc_indices"])
    assert torch.equal(B32.values, B64.values.to(torch.float32))


@pytest.mark.parametrize("name", ["fixtures_diffusion3d_p1_lognormal.npz", "fixtures_diffusion3d_p1_lognormal_24.npz"])
def test_cg_refine_assembled_operator(golden, name):
    """FP64 residuals by one SpMV pass on the assembled FP64 matrix, FP32
    filtered copy for the inner solves: within 1e-8 of the reference's tight
    solution; the CSR residual kernel equals SciPy's b - A x."""
    from fealpy_amd.solver.refine import _residual
    g = golden(name)
    _, space, form, F = _problem(g)
    bc = fa.DirichletBC(space, gd=U3)
    A64 = form.assembly()
    n = int(g["gdof"])
    bd = torch.as_tensor(g["is_bd"].astype(np.uint8), device=DEV)
    xv = np.random.default_rng(4).standard_normal(n)
    r, nr2 = _residual(A64, torch.as_tensor(g["F"], device=DEV), torch.as_tensor(xv, device=DEV), bd, torch.float64)
    A = sp.csr_matrix((g["A_data"], g["A_indices"], g["A_indptr"]), shape=(n, n))
    ref = g["F"] - A @ xv
    ref[g["is_bd"]] = 0.0
    np.testing.assert_allclose(_np(r), ref, rtol=0, atol=1e-12 * np.abs(ref).max() * 16)
    assert abs(float(nr2) - ref @ ref) <= 1e-12 * (ref @ ref)
    A32 = bc.apply_matrix(A64, dtype=torch.float32)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32).setup(A32)
    xs, info = cg_refine(A64, F, A32, mg, bc, rtol=1e-10, returninfo=True)
    err = np.linalg.norm(_np(xs) - g["x"]) / np.linalg.norm(g["x"])
    assert err <= 1e-8, (err, info)
