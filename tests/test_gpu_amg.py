"""Smoothed-aggregation AMG (fealpy_amd.solver.AlgebraicMG) as the CG
preconditioner, checked against direct solves of the oracle's system:

- the aggregation: every row with a strong connection in exactly one
  aggregate, aggregate numbers dense, roots pairwise more than two strong
  hops apart (checked on the host from the returned map), and the same map
  on a second run;
- P has one row per fine DOF, Dirichlet rows empty, and the hierarchy's
  Galerkin operators equal SciPy's R A P;
- MG-PCG solutions within 1e-8 of SciPy's direct solve on a from_box mesh
  and on a mesh with cells removed (threshold, so no geometric hierarchy),
  in FP64 and with the FP32 cycle; with iteration counts that stay low as
  the mesh is refined."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

import fealpy_amd as fa
from fealpy_amd import _native as nat
from fealpy_amd.solver import AlgebraicMG, GeometricMG

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")
PI = np.pi


def _np(t):
    return t.detach().cpu().numpy()


def _sp(A):
    return sp.csr_matrix((_np(A.values).astype(np.float64), _np(A.col), _np(A.crow)), shape=A.shape)


def _system(mesh, p=1, seed=3):
    space = fa.LagrangeFESpace(mesh, p)
    k = torch.as_tensor(np.exp(np.random.default_rng(seed).standard_normal(mesh.number_of_cells())), device=DEV)
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=p + 1)).assembly()
    F = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(1.0, q=p + 1)).assembly()
    bc = fa.DirichletBC(space, gd=0.0)
    A2, F2 = bc.apply(A, F)
    return space, bc, A2, F2


def test_aggregation_properties():
    mesh = fa.TetrahedronMesh.from_box(nx=10, ny=9, nz=8, device=DEV)
    space, bc, A, F = _system(mesh)
    L = nat.lib()
    n = A.shape[0]

    def run():
        agg = torch.empty(n, dtype=torch.int32, device=DEV)
        ws = nat.workspace(L.fa_amg_aggregate_work_bytes(n, A.nnz), DEV)
        na, rounds = nat.ctypes.c_int64(0), nat.ctypes.c_int32(0)
        nat.check(L.fa_amg_aggregate(nat.FA_F64, n, nat.ptr(A.crow), nat.ptr(A.col), nat.ptr(A.values), 0.08,
                                     nat.ptr(agg), nat.ctypes.byref(na), nat.ctypes.byref(rounds), nat.ptr(ws),
                                     ws.numel(), nat.stream()), "fa_amg_aggregate")
        return _np(agg), int(na.value), int(rounds.value)
    agg, na, rounds = run()
    agg2, _, _ = run()
    assert np.array_equal(agg, agg2)
    M = _sp(A)
    d = np.abs(M.diagonal())
    C = M.tocoo()
    strong = (C.row != C.col) & (np.abs(C.data) >= 0.08 * np.sqrt(d[C.row] * d[C.col]))
    S = sp.csr_matrix((np.ones(strong.sum()), (C.row[strong], C.col[strong])), shape=M.shape)
    has_strong = np.diff(S.indptr) > 0
    assert np.array_equal(agg >= 0, has_strong)                      # Dirichlet rows in no aggregate
    assert set(np.unique(agg[agg >= 0])) == set(range(na))            # dense numbering
    assert 1 <= rounds < 50 and na < 0.5 * has_strong.sum()
    # every aggregate is connected to its members by strong paths of length <= 2
    S2 = (S + S @ S + sp.identity(n)).tocsr()
    for a in range(0, na, max(1, na // 50)):
        members = np.nonzero(agg == a)[0]
        reach = S2[members[0]].indices
        assert len(members) >= 1
        assert np.all(np.isin(members, S2[members].indices))
        assert reach.size > 0


def test_galerkin_operators_equal_scipy():
    mesh = fa.TetrahedronMesh.from_box(nx=14, ny=14, nz=14, device=DEV)
    space, bc, A, F = _system(mesh)
    mg = AlgebraicMG(dtype=torch.float64).setup(A)
    assert len(mg.levels) >= 3 and mg.levels[-1].n <= 64
    for f, c in zip(mg.levels[:-1], mg.levels[1:]):
        P = _sp(f.P)
        assert P.shape == (f.n, c.n)
        ref = (P.T @ _sp(f.A) @ P).tocsr()
        got = _sp(c.A)
        diff = abs(got - ref).max() if got.nnz else 0.0
        assert diff <= 1e-12 * abs(ref).max()
    P0 = _sp(mg.levels[0].P)
    bd = _np(bc.is_boundary_dof)
    assert np.all(np.diff(P0.indptr)[bd] == 0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_amg_pcg_from_box_and_threshold(dtype):
    for mesh in (fa.TetrahedronMesh.from_box(nx=14, ny=14, nz=14, device=DEV),
                 fa.TetrahedronMesh.from_box(nx=14, ny=14, nz=14, device=DEV,
                                             threshold=lambda p: (p[..., 0] > 0.5) & (p[..., 1] > 0.5))):
        space, bc, A, F = _system(mesh)
        mg = AlgebraicMG(dtype=dtype)
        x, info = fa.cg(A, F, M=mg, atol=0.0, rtol=1e-10, maxit=300, returninfo=True)
        ref = spla.spsolve(_sp(A).tocsc(), _np(F))
        err = np.linalg.norm(_np(x) - ref) / np.linalg.norm(ref)
        assert err <= 1e-8, (err, info)
        assert info["niter"] <= 60, info


def test_amg_iterations_stay_bounded():
    its = []
    for n in (8, 16, 24):
        mesh = fa.TetrahedronMesh.from_box(nx=n, ny=n, nz=n, device=DEV)
        space, bc, A, F = _system(mesh)
        x, info = fa.cg(A, F, M=AlgebraicMG(dtype=torch.float32), atol=0.0, rtol=1e-8, maxit=300, returninfo=True)
        its.append(info["niter"])
    assert its[-1] <= 2 * its[0] + 10, its


def test_amg_2d_p2_and_geometric_refusal():
    mesh = fa.TriangleMesh.from_box([0, 1, 0, 1], nx=30, ny=30, device=DEV,
                                    threshold=lambda p: (p[..., 0] - 0.5) ** 2 + (p[..., 1] - 0.5) ** 2 < 0.04)
    space, bc, A, F = _system(mesh, p=2)
    with pytest.raises(ValueError):
        GeometricMG(space, bc.is_boundary_dof)
    x, info = fa.cg(A, F, M=AlgebraicMG(dtype=torch.float32), atol=0.0, rtol=1e-10, maxit=300, returninfo=True)
    ref = spla.spsolve(_sp(A).tocsc(), _np(F))
    assert np.linalg.norm(_np(x) - ref) / np.linalg.norm(ref) <= 1e-8, info
