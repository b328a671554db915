"""High-order element matrices on the matrix cores against the FP64 element
kernel (itself pinned to the reference fixtures at 1e-12): diffusion with
k(x) = 1 + 0.5 sin(2 pi x) cos(2 pi y) z per quadrature point plus mass 1,
degree-2p quadrature.  Bars: FP32 MFMA within 2e-6 of the largest entry,
BF16 inputs with FP32 accumulation within 2e-2 (BF16 keeps 8 mantissa bits)."""
import numpy as np
import pytest
import torch

import fealpy_amd as fa
from fealpy_amd.fem.integrators import GroupIntegrator

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")
PI = np.pi


@fa.cartesian
def kx(p):
    return 1 + 0.5 * torch.sin(2 * PI * p[..., 0]) * torch.cos(2 * PI * p[..., 1]) * p[..., 2]


@pytest.mark.parametrize("p,q", [(2, 4), (3, 5), (4, 6)])
@pytest.mark.parametrize("precision,tol", [("fp32", 2e-6), ("bf16", 2e-2)])
def test_mfma_elements_match_fp64(p, q, precision, tol):
    mesh = fa.TetrahedronMesh.from_box([0, 1, 0, 1, 0, 1], nx=6, ny=5, nz=4, device=DEV)   # 720 cells, 2 blocks
    space = fa.LagrangeFESpace(mesh, p)
    ref = GroupIntegrator(fa.ScalarDiffusionIntegrator(kx, q=q), fa.ScalarMassIntegrator(1.0, q=q)).assembly(space)
    m = f"mfma_{precision}"
    grp = GroupIntegrator(fa.ScalarDiffusionIntegrator(kx, q=q, method=m), fa.ScalarMassIntegrator(1.0, q=q, method=m))
    got = grp.assembly(space, dtype=torch.float32)
    assert got.dtype == torch.float32 and got.shape == ref.shape
    assert torch.equal(grp.assembly(space), got.double())        # default: the same values as FP64
    err = float((got.double() - ref).abs().max()) / float(ref.abs().max())
    assert err < tol, err


def test_mfma_assembled_matrix_and_cell_coefficient():
    """Per-cell coefficient, diffusion only, assembled through BilinearForm."""
    mesh = fa.TetrahedronMesh.from_box([0, 1, 0, 1, 0, 1], nx=4, ny=4, nz=4, device=DEV)
    space = fa.LagrangeFESpace(mesh, 3)
    kc = torch.as_tensor(np.exp(np.random.default_rng(2).standard_normal(mesh.number_of_cells())), device=DEV)
    A64 = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(kc, q=5)).assembly()
    A32 = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(kc, q=5, method="mfma_fp32")).assembly(
        dtype=torch.float32)
    assert torch.equal(A64.col, A32.col) and A32.values.dtype == torch.float32
    err = float((A32.values.double() - A64.values).abs().max()) / float(A64.values.abs().max())
    assert err < 2e-6, err
