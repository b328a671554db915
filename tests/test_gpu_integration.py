"""The reference-side binding (fealpy_amd/integration/fealpy_backend.py, the
file a fealpy maintainer would add as fealpy/backend/mi355x.py): installed on
a backend class the way it would be installed on PyTorchBackend, then called
through the class like fealpy's CSRTensor.matmul calls bm.csr_spmm.  Checked
against SciPy's csr_matvec / csr_matvecs (the numpy backend's hook) on 1-D
and 2-D operands, FP64 and FP32."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import fealpy_amd as fa
from fealpy_amd.integration import fealpy_backend as fb

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


class _TorchBackendStandIn:
    """Stand-in for fealpy.backend.pytorch_backend.PyTorchBackend, which
    defines csr_spmm itself (so attach_methods would not replace it)."""

    @staticmethod
    def csr_spmm(crow, col, values, shape, other):
        raise AssertionError("the binding was not installed")


def _matrix():
    mesh = fa.TetrahedronMesh.from_box(nx=7, ny=6, nz=5, device=DEV)
    space = fa.LagrangeFESpace(mesh, 2)
    g = torch.Generator().manual_seed(2)
    k = torch.rand(mesh.number_of_cells(), generator=g, dtype=torch.float64).to(DEV) + 0.5
    return fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=4),
                                                 fa.ScalarMassIntegrator(1.0, q=4)).assembly()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_registered_csr_spmm_matches_scipy(dtype):
    fb.register(_TorchBackendStandIn)
    A = _matrix()
    crow, col = A.crow.long(), A.col.long()        # the reference's int64 crow / int indices
    vals = A.values.to(dtype)
    S = sp.csr_matrix((vals.cpu().numpy(), col.cpu().numpy(), crow.cpu().numpy()), shape=A.shape)
    g = np.random.default_rng(7)
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    for shape in ((A.shape[1],), (A.shape[1], 5), (A.shape[1], 1)):
        xh = g.standard_normal(shape).astype(np.float64 if dtype == torch.float64 else np.float32)
        y = _TorchBackendStandIn.csr_spmm(crow, col, vals, A.shape, torch.as_tensor(xh, device=DEV))
        ref = S @ xh
        assert tuple(y.shape) == ref.shape and y.dtype == dtype
        np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=0, atol=tol * np.abs(ref).max())
    with pytest.raises(ValueError):
        _TorchBackendStandIn.csr_spmm(crow, col, vals, A.shape, torch.zeros((A.shape[1], 2, 2), dtype=dtype,
                                                                            device=DEV))
    with pytest.raises(NotImplementedError):
        _TorchBackendStandIn.csr_spmm(crow, col, vals[None], A.shape, torch.zeros(A.shape[1], dtype=dtype,
                                                                                  device=DEV))


def test_csrtensor_matmul_2d_matches_scipy():
    A = _matrix()
    X = torch.rand((A.shape[1], 9), dtype=torch.float64, device=DEV)
    Y = A @ X
    ref = A.to_scipy() @ X.cpu().numpy()
    np.testing.assert_allclose(Y.cpu().numpy(), ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    for k in range(9):                              # each column equals the 1-D product
        np.testing.assert_allclose(Y[:, k].cpu().numpy(), (A @ X[:, k].contiguous()).cpu().numpy(),
                                   rtol=0, atol=1e-14 * np.abs(ref).max())
