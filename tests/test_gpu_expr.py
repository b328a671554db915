"""Coefficients and sources evaluated inside the element kernels
(FA_COEF_SEP) against the reference: the same fixtures the torch-callable
path is checked on (values written by running the reference's numpy
backend), at the same 1e-12 relative bar, on every kernel that reads the
table -- the generic element matrices and vectors (P2-P4), the fused P1
kernels and the two-pass P1 passes."""
import numpy as np
import pytest
import torch

import fealpy_amd as fa
from fealpy_amd.expr import cell, cos, sin, x, y, z
from fealpy_amd.mesh import TetrahedronMesh, TriangleMesh

pytestmark = pytest.mark.gpu
PI = np.pi
DEV = torch.device("cuda")


def _np(t):
    return t.detach().cpu().numpy()


def _close(got, ref, rtol=1e-12):
    ref = np.asarray(ref)
    np.testing.assert_allclose(_np(got), ref, rtol=0, atol=rtol * np.abs(ref).max())


U3 = sin(PI * x) * sin(PI * y) * sin(PI * z)


def _k_sep():
    return 1 + 0.5 * sin(2 * PI * x) * cos(2 * PI * y) * z


def _f_highorder_sep():
    """-div(k grad u) + u for u = sin sin sin, k = 1 + 0.5 sin(2 pi x) cos(2 pi y) z."""
    k = _k_sep()
    dk = (PI * cos(2 * PI * x) * cos(2 * PI * y) * z, -PI * sin(2 * PI * x) * sin(2 * PI * y) * z,
          0.5 * sin(2 * PI * x) * cos(2 * PI * y))
    du = (PI * cos(PI * x) * sin(PI * y) * sin(PI * z), PI * sin(PI * x) * cos(PI * y) * sin(PI * z),
          PI * sin(PI * x) * sin(PI * y) * cos(PI * z))
    return 3 * PI ** 2 * k * U3 - (dk[0] * du[0] + dk[1] * du[1] + dk[2] * du[2]) + U3


@pytest.mark.parametrize("p", [3, 4])
def test_highorder_separable_coefficient_and_source(golden, p):
    F = golden(f"fixtures_highorder3d_p{p}.npz")
    n, q = int(F["n"]), int(F["q"])
    mesh = TetrahedronMesh.from_box(nx=n, ny=n, nz=n, device=DEV)
    space = fa.LagrangeFESpace(mesh, p)
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(_k_sep(), q=q),
                                              fa.ScalarMassIntegrator(1.0, q=q)).assembly()
    np.testing.assert_array_equal(_np(A.col), F["A_indices"])
    _close(A.values, F["A_data"])
    b = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(_f_highorder_sep(), q=q)).assembly()
    _close(b, F["F"])


@pytest.mark.parametrize("path", ["fused", "two_pass", "box"])
def test_lognormal_source_in_kernel(golden, path, monkeypatch):
    """f_c(x) = 3 pi^2 k_c u(x) with the per-cell field read in the kernel:
    fused P1 kernel, the two-pass passes (forced on the small mesh) and, for
    matrix and vector, the structured from_box kernels."""
    from fealpy_amd.fem import integrators
    if path == "two_pass":
        monkeypatch.setattr(integrators, "TWO_PASS_MIN_CELLS", 0)
    if path != "box":
        monkeypatch.setattr(integrators, "BOX_ASSEMBLY", False)
    F = golden("fixtures_diffusion3d_p1_lognormal.npz")
    n = int(F["n"])
    mesh = TetrahedronMesh.from_box(nx=n, ny=n, nz=n, device=DEV)
    space = fa.LagrangeFESpace(mesh, 1)
    k = torch.as_tensor(F["k"], device=DEV)
    f = 3 * PI ** 2 * cell(k) * U3
    b = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(f, q=2)).assembly()
    _close(b, F["F"])
    # the coefficient as an expression through the P1 matrix kernels
    kx = _k_sep()
    A1 = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(kx, q=3),
                                               fa.ScalarMassIntegrator(cell(k) * (1 + x), q=3)).assembly()
    A2 = fa.BilinearForm(space).add_integrator(
        fa.ScalarDiffusionIntegrator(fa.cartesian(lambda pt: kx(pt)), q=3),
        fa.ScalarMassIntegrator(fa.cartesian(lambda pt: k[:, None] * (1 + pt[..., 0])), q=3)).assembly()
    assert torch.equal(A1.col, A2.col)
    _close(A1.values, _np(A2.values), 1e-14)


def test_fourier_source_2d_p2(golden):
    F = golden("fixtures_poisson2d_p2_fourier.npz")
    n = int(F["n"])
    mesh = TriangleMesh.from_box([0, 1, 0, 1], n, n, device=DEV)
    space = fa.LagrangeFESpace(mesh, 2)
    f = sum(float(F["a"][m]) * sin(float(F["km"][m]) * PI * x) * sin(float(F["lm"][m]) * PI * y) for m in range(8))
    b = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(f)).assembly()
    _close(b, F["F"])


def test_expression_as_dirichlet_data_and_error():
    """The same object serves as cartesian Dirichlet data and in mesh.error."""
    mesh = TetrahedronMesh.from_box(nx=6, ny=6, nz=6, device=DEV)
    space = fa.LagrangeFESpace(mesh, 1)
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(q=2)).assembly()
    b = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(3 * PI ** 2 * U3, q=2)).assembly()
    A2, b2 = fa.DirichletBC(space, gd=U3).apply(A, b)
    xh = fa.cg(A2, b2, atol=1e-14, rtol=1e-13)
    e1 = float(mesh.error(U3, space.function(xh)))
    u_t = fa.cartesian(lambda pt: torch.sin(PI * pt).prod(-1))
    e2 = float(mesh.error(u_t, space.function(xh)))
    assert abs(e1 - e2) <= 1e-12 * e2 and 0 < e1 < 0.1


def test_h1_error_variant():
    """mesh.error(grad u, uh.grad_value), the reference's H1 semi-norm error:
    zero for a function the space holds (P2 interpolant of a quadratic),
    and O(h^p) under refinement for the P1 and P2 solutions of -lap u = f."""
    mesh = TetrahedronMesh.from_box(nx=3, ny=3, nz=3, device=DEV)
    space = fa.LagrangeFESpace(mesh, 2)
    quad = fa.cartesian(lambda p: p[..., 0] ** 2 + 2 * p[..., 1] * p[..., 2] - p[..., 2])
    gquad = fa.cartesian(lambda p: torch.stack([2 * p[..., 0], 2 * p[..., 2], 2 * p[..., 1] - 1], dim=-1))
    uh = space.interpolate(quad)
    assert float(mesh.error(gquad, uh.grad_value, q=4)) < 1e-12
    gu = fa.cartesian(lambda p: PI * torch.stack([torch.cos(PI * p[..., 0]) * torch.sin(PI * p[..., 1]) *
                                                  torch.sin(PI * p[..., 2]),
                                                  torch.sin(PI * p[..., 0]) * torch.cos(PI * p[..., 1]) *
                                                  torch.sin(PI * p[..., 2]),
                                                  torch.sin(PI * p[..., 0]) * torch.sin(PI * p[..., 1]) *
                                                  torch.cos(PI * p[..., 2])], dim=-1))
    for p in (1, 2):
        errs = []
        for n in (4, 8):
            m = TetrahedronMesh.from_box(nx=n, ny=n, nz=n, device=DEV)
            sp = fa.LagrangeFESpace(m, p)
            A = fa.BilinearForm(sp).add_integrator(fa.ScalarDiffusionIntegrator(q=p + 2)).assembly()
            b = fa.LinearForm(sp).add_integrator(fa.ScalarSourceIntegrator(3 * PI ** 2 * U3, q=p + 2)).assembly()
            A2, b2 = fa.DirichletBC(sp, gd=U3).apply(A, b)
            xh = fa.cg(A2, b2, atol=1e-14, rtol=1e-13, maxit=5000)
            errs.append(float(m.error(gu, sp.function(xh).grad_value, q=4)))
        rate = np.log2(errs[0] / errs[1])
        assert p - 0.25 < rate < p + 0.5, (p, errs, rate)


@pytest.mark.parametrize("p", [1, 2])
def test_every_factor_kind_matches_torch(p, monkeypatch):
    """sin, cos (through sin(pi t) / cos(pi t) in the kernel), exp, affine and
    whole powers, with large arguments and a cell field: the in-kernel values
    against the same expression evaluated by torch at the quadrature points."""
    from fealpy_amd.expr import exp
    from fealpy_amd.fem import integrators
    monkeypatch.setattr(integrators, "TWO_PASS_MIN_CELLS", 0)
    monkeypatch.setattr(integrators, "BOX_ASSEMBLY", False)
    mesh = TetrahedronMesh.from_box([-3, 5, 0, 2, -1, 1], nx=5, ny=4, nz=3, device=DEV)
    space = fa.LagrangeFESpace(mesh, p)
    k = torch.rand(mesh.number_of_cells(), dtype=torch.float64, device=DEV) + 0.5
    f = (2.5 * exp(1.7 * x - 0.3) * cos(7.1 * y + 0.2) + cell(k) * (x - 0.5) ** 3 * sin(40.0 * z + 1.0)
         - exp(-2.0 * z) * (3 * y + 1) + sin(123.4 * x) * cos(PI * z))
    g = fa.cartesian(lambda pt: f(pt))                    # the same expression, evaluated with torch
    b1 = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(f, q=4)).assembly()
    b2 = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(
        fa.cartesian(lambda pt: f(pt)), q=4)).assembly()
    _close(b1, _np(b2), 1e-13)
    kk = 1.5 + sin(3 * x) * exp(0.1 * y) * z ** 2
    A1 = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(kk, q=4)).assembly()
    A2 = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(
        fa.cartesian(lambda pt: kk(pt)), q=4)).assembly()
    _close(A1.values, _np(A2.values), 1e-13)
    del g


def test_expr_at_device_points_matches_numpy():
    """A field-free expression at device points (one fa_sep_points launch)
    against numpy's FP64 functions on the same points, for every factor kind;
    the points' leading shape is kept."""
    from fealpy_amd.expr import exp
    rng = np.random.default_rng(7)
    pts = rng.uniform(-1.0, 2.0, size=(5, 7, 3))
    f = (3.0 * sin(PI * x) * cos(2.0 * y + 0.3) - exp(0.5 * z) * (2.0 * x - 1.0) + y ** 3 * z
         + 0.25 * cos(PI * (z - 0.5)) * sin(1.7 * x))
    X, Y, Z = pts[..., 0], pts[..., 1], pts[..., 2]
    ref = (3.0 * np.sin(PI * X) * np.cos(2.0 * Y + 0.3) - np.exp(0.5 * Z) * (2.0 * X - 1.0) + Y ** 3 * Z
           + 0.25 * np.cos(PI * (Z - 0.5)) * np.sin(1.7 * X))
    got = f(torch.tensor(pts, device=DEV))
    assert got.shape == (5, 7) and got.is_cuda
    _close(got, ref, rtol=1e-14)
    # the kernel ran (the result is bitwise the same on a second call, and the
    # CPU path agrees to rounding)
    assert torch.equal(got, f(torch.tensor(pts, device=DEV)))
    _close(f(torch.tensor(pts)), ref, rtol=1e-14)
    # 2D points, a single point
    p2 = rng.uniform(0, 1, size=(11, 2))
    g = sin(PI * x) * sin(PI * y)
    _close(g(torch.tensor(p2, device=DEV)), np.sin(PI * p2[:, 0]) * np.sin(PI * p2[:, 1]), rtol=1e-15)
    one = g(torch.tensor(p2[0], device=DEV))
    assert one.shape == () and abs(float(one) - np.sin(PI * p2[0, 0]) * np.sin(PI * p2[0, 1])) < 1e-15
