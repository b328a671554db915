"""fa_pcg_static (the PCG entry point that allocates nothing: its workspace is
a file-scope device array) against fa_pcg with a caller workspace, through
the C-ABI: bitwise the same x, iteration count and residual, on two calls in
a row with the second system smaller than the first, with and without the
multigrid preconditioner; sizes above the static capacity are refused."""
import ctypes

import pytest
import torch

import fealpy_amd as fa
from fealpy_amd import _native as nat
from fealpy_amd.sparse.sell import FaSell, sell_of

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def _system(n):
    mesh = fa.TetrahedronMesh.from_box(nx=n, ny=n, nz=n, device=DEV)
    space = fa.LagrangeFESpace(mesh, 1)
    g = torch.Generator().manual_seed(n)
    k = torch.exp(torch.randn(mesh.number_of_cells(), generator=g, dtype=torch.float64)).to(DEV)
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=2)).assembly()
    b = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(1.0, q=2)).assembly()
    bc = fa.DirichletBC(space, gd=0.0)
    A2, b2 = bc.apply(A, b)
    return space, bc.is_boundary_dof, A2, b2


def _static(A_desc, M_desc, b, rtol, check_every, maxit=400):
    L = nat.lib()
    x = torch.zeros_like(b)
    niter, res = ctypes.c_int64(0), ctypes.c_double(0.0)
    st = L.fa_pcg_static(nat.FA_F64, ctypes.byref(A_desc), M_desc, nat.ptr(b), nat.ptr(x), 0.0, float(rtol),
                         maxit, check_every, ctypes.byref(niter), ctypes.byref(res), nat.stream())
    nat.check(st, "fa_pcg_static")
    return x, int(niter.value), float(res.value)


@pytest.mark.parametrize("precond", [True, False])
def test_static_pcg_equals_workspace_pcg(precond):
    for n in (24, 12):                                   # second call on a smaller system
        space, bd, A, b = _system(n)
        if precond:
            mg = fa.GeometricMG(space, bd, dtype=torch.float32)
            x_ref, info = mg.pcg(A, b, atol=0.0, rtol=1e-10, maxit=400, check_every=4)
            d = mg.levels[0].As.desc(torch.float64)
            x, it, res = _static(d, ctypes.byref(mg._desc), b, 1e-10, 4)
        else:
            x_ref, info = fa.cg(A, b, atol=0.0, rtol=1e-10, maxit=400, returninfo=True)
            d = sell_of(A).desc(torch.float64)
            x, it, res = _static(d, None, b, 1e-10, 16)
        torch.cuda.synchronize()
        assert it == info["niter"] and res == info["residual"], (n, it, info)
        assert torch.equal(x, x_ref), n


def test_static_pcg_refuses_oversized_systems():
    """The size check happens on the host, before any GPU work."""
    L = nat.lib()
    big = FaSell(262145, 262145, None, None, None)
    niter, res = ctypes.c_int64(0), ctypes.c_double(0.0)
    st = L.fa_pcg_static(nat.FA_F64, ctypes.byref(big), None, None, None, 0.0, 1e-8, 10, 4,
                         ctypes.byref(niter), ctypes.byref(res), nat.stream())
    assert st == nat.FA_EINVAL
    assert b"static workspace" in L.fa_last_error()
