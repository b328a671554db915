"""The pattern-preserving Dirichlet mode (DirichletBC.apply(...,
keep_pattern=True), fa_dirichlet_apply_keep) against the reference-compatible
filtered mode, which the reference's fixtures pin:

- the matrix keeps A's row pointer and column tensors; as a dense matrix it
  equals the filtered matrix exactly (unit rows on the boundary, boundary
  columns zero); the right-hand side is bitwise the filtered one;
- in place (new_values = values) equals out of place;
- multigrid-preconditioned CG on it reaches the filtered system's solution
  with the same iteration count, and a second boundary condition on the same
  pattern reuses the hierarchy's symbolic products."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import fealpy_amd as fa
from fealpy_amd import _native as nat
from fealpy_amd.solver import GeometricMG

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def _dense(A):
    return sp.csr_matrix((A.values.double().cpu().numpy(), A.col.cpu().numpy(), A.crow.cpu().numpy()),
                         shape=A.shape).toarray()


@pytest.mark.parametrize("case", ["tet_p1", "tri_p2", "tet_p2_f32"])
def test_keep_equals_filtered(case):
    if case == "tri_p2":
        mesh, p, dt = fa.TriangleMesh.from_box([0, 1, 0, 1], nx=8, ny=7, device=DEV), 2, torch.float64
    else:
        mesh = fa.TetrahedronMesh.from_box(nx=5, ny=4, nz=6, device=DEV)
        p, dt = (1, torch.float64) if case == "tet_p1" else (2, torch.float32)
    space = fa.LagrangeFESpace(mesh, p)
    k = torch.as_tensor(np.exp(np.random.default_rng(5).standard_normal(mesh.number_of_cells())), device=DEV)
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=p + 2)).assembly(dtype=dt)
    F = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(1.0, q=p + 2)).assembly(dtype=dt)
    gd = fa.cartesian(lambda x: 1.0 + x[..., 0] * x[..., 1] - x[..., -1])
    bc = fa.DirichletBC(space, gd=gd)
    Af, Ff = bc.apply(A, F)
    Ak, Fk = bc.apply(A, F, keep_pattern=True)
    assert Ak.crow is A.crow and Ak.col is A.col and Ak.values.dtype == dt
    assert torch.equal(Fk, Ff)
    assert np.array_equal(_dense(Ak), _dense(Af))
    bd = bc.is_boundary_dof.cpu().numpy()
    D = _dense(Ak)
    assert np.array_equal(D[bd], np.eye(D.shape[0])[bd]) and np.all(D[~bd][:, bd] == 0)
    # in place: new_values = values
    vals = A.values.clone()
    g = bc._boundary_values(F, None, gd, dt)
    out = torch.empty_like(Fk)
    nat.check(nat.lib().fa_dirichlet_apply_keep(nat.dtype_code(dt), A.shape[0], nat.ptr(A.crow), nat.ptr(A.col),
                                                nat.ptr(vals), nat.ptr(bc._bd_u8), nat.ptr(vals), nat.ptr(g),
                                                nat.ptr(F.to(dt).contiguous()), nat.ptr(out), nat.stream()),
              "fa_dirichlet_apply_keep")
    assert torch.equal(vals, Ak.values) and torch.equal(out, Fk)


def test_keep_pattern_mg_solve_and_reuse():
    mesh = fa.TetrahedronMesh.from_box(nx=12, ny=12, nz=12, device=DEV)
    space = fa.LagrangeFESpace(mesh, 1)
    k = torch.as_tensor(np.exp(np.random.default_rng(7).standard_normal(mesh.number_of_cells())), device=DEV)
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=2)).assembly()
    F = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(1.0, q=2)).assembly()
    mgk = None
    for gd in (0.0, fa.cartesian(lambda x: torch.sin(3 * x[..., 0]) + x[..., 2])):
        bc = fa.DirichletBC(space, gd=gd)
        Af, Ff = bc.apply(A, F)
        Ak, Fk = bc.apply(A, F, keep_pattern=True)
        mgf = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32)
        if mgk is None:
            mgk = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32)
            key = None
        xf, inf = fa.cg(Af, Ff, M=mgf, atol=0.0, rtol=1e-10, maxit=200, returninfo=True)
        xk, ink = fa.cg(Ak, Fk, M=mgk, atol=0.0, rtol=1e-10, maxit=200, returninfo=True)
        if key is None:
            key = mgk._pattern_key
        assert mgk._pattern_key == key                 # the second condition kept the symbolic products
        assert abs(inf["niter"] - ink["niter"]) <= 1, (inf, ink)
        err = float(torch.linalg.norm(xk - xf) / torch.linalg.norm(xf))
        assert err <= 1e-9, err
