"""Edge semantics of the drop-in API checked against the oracle: a 2-D
right-hand side through DirichletBC (one column at a time, as the
reference's A.matmul of a 2-D uh), cg with b = 0 and with maxit = 0, and
the value type of the MFMA element path."""
import numpy as np
import pytest
import torch

import fealpy_amd as fa
from oracle import fem as ofem
from oracle.mesh import tet_from_box

pytestmark = pytest.mark.gpu
PI = np.pi
DEV = torch.device("cuda")


def test_dirichlet_2d_rhs_matches_oracle_per_column():
    n = 5
    mesh = fa.TetrahedronMesh.from_box(nx=n, ny=n, nz=n, device=DEV)
    space = fa.LagrangeFESpace(mesh, 2)
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(q=4)).assembly()
    g = np.random.default_rng(4)
    F = torch.as_tensor(g.standard_normal((space.number_of_global_dofs(), 3)), device=DEV)
    u = fa.cartesian(lambda p: torch.sin(PI * p[..., 0]) * p[..., 1] + p[..., 2])
    bc = fa.DirichletBC(space, gd=u)
    A2, F2 = bc.apply(A, F)
    assert F2.shape == F.shape
    o = tet_from_box(nx=n, ny=n, nz=n)
    c2d = o.cell_to_ipoint(2)
    Ao = ofem.assemble_matrix(c2d, [ofem.diffusion_element(o, 2, None, 4)], o.number_of_global_ipoints(2))
    uo = ofem.cartesian(lambda p: np.sin(PI * p[..., 0]) * p[..., 1] + p[..., 2])
    Fh = F.cpu().numpy()
    for j in range(3):
        Ao2, fo = ofem.dirichlet_apply(Ao, Fh[:, j].copy(), o.is_boundary_dof(2), o.interpolation_points(2), uo)
        np.testing.assert_allclose(F2[:, j].cpu().numpy(), fo, rtol=0, atol=1e-12 * np.abs(fo).max())
    np.testing.assert_array_equal(A2.col.cpu().numpy(), Ao2.indices)
    # apply_vector alone gives the same columns
    F3 = bc.apply_vector(F, A)
    assert torch.equal(F3, F2)


def test_cg_zero_rhs_and_maxit_zero():
    mesh = fa.TetrahedronMesh.from_box(nx=6, ny=6, nz=6, device=DEV)
    space = fa.LagrangeFESpace(mesh, 1)
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(q=2)).assembly()
    A2, _ = fa.DirichletBC(space, gd=0.0).apply(A, torch.zeros(A.shape[0], dtype=torch.float64, device=DEV))
    b0 = torch.zeros(A.shape[0], dtype=torch.float64, device=DEV)
    bd = fa.DirichletBC(space, gd=0.0).is_boundary_dof
    for M in (None, fa.GeometricMG(space, bd)):
        x, info = fa.cg(A2, b0, M=M, returninfo=True)
        assert info["residual"] == 0.0 and info["niter"] == 0 and not torch.any(x)
    b = torch.rand(A.shape[0], dtype=torch.float64, device=DEV)
    b[bd] = 0
    for M in (None, fa.GeometricMG(space, bd)):
        x, info = fa.cg(A2, b, M=M, maxit=0, returninfo=True)
        assert info["niter"] == 1            # the reference stops after its first update
    Ao = A2.to_scipy()
    xo, io = ofem.cg(Ao, b.cpu().numpy(), maxit=0)
    assert io["niter"] == 1
    x, info = fa.cg(A2, b, maxit=0, returninfo=True)
    np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=0, atol=1e-12 * np.abs(xo).max())


def test_mfma_path_defaults_to_fp64_values():
    mesh = fa.TetrahedronMesh.from_box(nx=3, ny=3, nz=3, device=DEV)
    space = fa.LagrangeFESpace(mesh, 3)
    it = fa.ScalarDiffusionIntegrator(1.0, q=5, method='mfma_fp32')
    assert it.assembly(space).dtype == torch.float64
    assert it.assembly(space, dtype=torch.float32).dtype == torch.float32
    A = fa.BilinearForm(space).add_integrator(it).assembly()
    assert A.values.dtype == torch.float64


def test_native_stream_is_torch_current_stream():
    """The library launches on the calling thread's current torch stream,
    inside and outside a stream context."""
    import torch
    from fealpy_amd import _native as nat
    def handle():
        return nat.stream().value or 0                  # the null (default) stream is 0
    assert handle() == torch.cuda.current_stream().cuda_stream
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        assert handle() == s.cuda_stream != 0
    assert handle() == torch.cuda.current_stream().cuda_stream
