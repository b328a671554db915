"""Parity of the HIP path with the reference (fixtures written by running it)
and with the oracle.  Tolerances: CSR patterns bit-exact (integer work);
FP64 values 1e-12 relative to the largest entry; FP64 solutions 1e-10."""
import numpy as np
import pytest
import torch

import fealpy_amd as fa
from fealpy_amd.functionspace import LagrangeFESpace
from fealpy_amd.mesh import TetrahedronMesh, TriangleMesh
from oracle import fem as ofem
from oracle.mesh import tet_from_box

pytestmark = pytest.mark.gpu
PI = np.pi
DEV = torch.device("cuda")


def _np(t):
    return t.detach().cpu().numpy()


def _assert_csr(A, F, prefix, rtol=1e-12):
    np.testing.assert_array_equal(_np(A.crow).astype(np.int64), F[prefix + "_indptr"])
    np.testing.assert_array_equal(_np(A.col), F[prefix + "_indices"])
    ref = F[prefix + "_data"]
    np.testing.assert_allclose(_np(A.values), ref, rtol=0, atol=rtol * np.abs(ref).max())


def _assert_vec(v, ref, rtol=1e-12):
    np.testing.assert_allclose(_np(v), ref, rtol=0, atol=rtol * np.abs(ref).max())


def u_sin2(p):
    return torch.sin(PI * p[..., 0]) * torch.sin(PI * p[..., 1])


def u_sin3(p):
    return torch.sin(PI * p[..., 0]) * torch.sin(PI * p[..., 1]) * torch.sin(PI * p[..., 2])


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_tet_patterns_and_values(golden, p):
    F = golden("fixtures_patterns.npz")
    mesh = TetrahedronMesh.from_box(nx=4, ny=4, nz=4, device=DEV)
    space = LagrangeFESpace(mesh, p)
    np.testing.assert_array_equal(_np(space.cell_to_dof()), F[f"tet4_p{p}_cell2dof"])
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(q=p + 1)).assembly()
    _assert_csr(A, F, f"tet4_p{p}")


def test_poisson2d_p1_full(golden):
    F = golden("fixtures_poisson2d_p1.npz")
    mesh = TriangleMesh.from_box([0, 1, 0, 1], 64, 64, device=DEV)
    space = LagrangeFESpace(mesh, 1)
    u = fa.cartesian(u_sin2)
    f = fa.cartesian(lambda p: 2 * PI ** 2 * u_sin2(p))
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(q=2)).assembly()
    _assert_csr(A, F, "A")
    b = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(f, q=2)).assembly()
    _assert_vec(b, F["F"])
    bc = fa.DirichletBC(space, gd=u)
    np.testing.assert_array_equal(_np(bc.is_boundary_dof), F["is_bd"])
    A2, b2 = bc.apply(A, b)
    _assert_csr(A2, F, "Abc")
    _assert_vec(b2, F["Fbc"])
    x, info = fa.cg(A2, b2, atol=1e-14, rtol=1e-14, maxit=20000, returninfo=True)
    _assert_vec(x, F["x"], 1e-10)
    x2, info2 = fa.cg(A2, b2, rtol=1e-10, returninfo=True)
    assert abs(info2["niter"] - int(F["cg_niter"])) <= 2, (info2, int(F["cg_niter"]))
    uh = space.function(x)
    err = float(mesh.error(u, uh))
    assert abs(err - float(F["l2_error"])) <= 1e-6 * float(F["l2_error"])


def test_diffusion3d_p1_lognormal(golden):
    F = golden("fixtures_diffusion3d_p1_lognormal.npz")
    n = int(F["n"])
    mesh = TetrahedronMesh.from_box(nx=n, ny=n, nz=n, device=DEV)
    space = LagrangeFESpace(mesh, 1)
    k = torch.as_tensor(F["k"], device=DEV)

    @fa.barycentric
    def f(bcs, index=None):
        return 3 * PI ** 2 * k[:, None] * u_sin3(mesh.bc_to_point(bcs))

    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=2)).assembly()
    _assert_csr(A, F, "A")
    b = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(f, q=2)).assembly()
    _assert_vec(b, F["F"])
    A2, b2 = fa.DirichletBC(space, gd=fa.cartesian(u_sin3)).apply(A, b)
    _assert_csr(A2, F, "Abc")
    _assert_vec(b2, F["Fbc"])
    x = fa.cg(A2, b2, atol=1e-14, rtol=1e-14, maxit=20000)
    _assert_vec(x, F["x"], 1e-10)
    err = float(mesh.error(fa.cartesian(u_sin3), space.function(x)))
    assert abs(err - float(F["l2_error"])) <= 1e-8 * float(F["l2_error"])


def _k_smooth(p):
    return 1 + 0.5 * torch.sin(2 * PI * p[..., 0]) * torch.cos(2 * PI * p[..., 1]) * p[..., 2]


def _f_highorder(pt):
    x, y, z = pt[..., 0], pt[..., 1], pt[..., 2]
    sx, sy, sz = torch.sin(PI * x), torch.sin(PI * y), torch.sin(PI * z)
    cx, cy, cz = torch.cos(PI * x), torch.cos(PI * y), torch.cos(PI * z)
    uu = sx * sy * sz
    gk = (PI * torch.cos(2 * PI * x) * torch.cos(2 * PI * y) * z, -PI * torch.sin(2 * PI * x) * torch.sin(2 * PI * y) * z,
          0.5 * torch.sin(2 * PI * x) * torch.cos(2 * PI * y))
    gu = (PI * cx * sy * sz, PI * sx * cy * sz, PI * sx * sy * cz)
    return 3 * PI ** 2 * _k_smooth(pt) * uu - sum(a * b for a, b in zip(gk, gu)) + uu


@pytest.mark.parametrize("p", [3, 4])
def test_highorder3d_diffusion_mass(golden, p):
    F = golden(f"fixtures_highorder3d_p{p}.npz")
    n, q = int(F["n"]), int(F["q"])
    mesh = TetrahedronMesh.from_box(nx=n, ny=n, nz=n, device=DEV)
    space = LagrangeFESpace(mesh, p)
    bform = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(fa.cartesian(_k_smooth), q=q),
                                                  fa.ScalarMassIntegrator(1.0, q=q))
    A = bform.assembly()
    _assert_csr(A, F, "A")
    b = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(fa.cartesian(_f_highorder), q=q)).assembly()
    _assert_vec(b, F["F"])
    A2, b2 = fa.DirichletBC(space, gd=fa.cartesian(u_sin3)).apply(A, b)
    _assert_csr(A2, F, "Abc")
    x = fa.cg(A2, b2, atol=1e-14, rtol=1e-14, maxit=20000)
    _assert_vec(x, F["x"], 1e-10)
    err = float(mesh.error(fa.cartesian(u_sin3), space.function(x), q=q))
    assert abs(err - float(F["l2_error"])) <= 1e-8 * float(F["l2_error"])


def test_poisson2d_p2_fourier(golden):
    F = golden("fixtures_poisson2d_p2_fourier.npz")
    n = int(F["n"])
    a, km, lm = (torch.as_tensor(F[k], dtype=torch.float64, device=DEV) for k in ("a", "km", "lm"))
    mesh = TriangleMesh.from_box([0, 1, 0, 1], n, n, device=DEV)
    space = LagrangeFESpace(mesh, 2)

    @fa.cartesian
    def f(p):
        return sum(a[m] * torch.sin(km[m] * PI * p[..., 0]) * torch.sin(lm[m] * PI * p[..., 1]) for m in range(8))

    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(torch.as_tensor(F["k"], device=DEV))
                                              ).assembly()
    _assert_csr(A, F, "A")
    b = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(f)).assembly()
    _assert_vec(b, F["F"])
    A2, b2 = fa.DirichletBC(space, gd=fa.cartesian(lambda p: torch.zeros(p.shape[:-1], dtype=p.dtype,
                                                                           device=p.device))).apply(A, b)
    _assert_csr(A2, F, "Abc")
    x = fa.cg(A2, b2, atol=1e-14, rtol=1e-14, maxit=20000)
    _assert_vec(x, F["x"], 1e-10)
    zero = fa.cartesian(lambda p: torch.zeros(p.shape[:-1], dtype=p.dtype, device=p.device))
    err = float(mesh.error(zero, space.function(x)))
    assert abs(err - float(F["l2_error"])) <= 1e-8 * float(F["l2_error"])


def test_assembly_and_cg_are_bitwise_reproducible():
    mesh = TetrahedronMesh.from_box(nx=12, ny=12, nz=12, device=DEV)
    space = LagrangeFESpace(mesh, 2)
    k = torch.exp(torch.randn(mesh.number_of_cells(), generator=torch.Generator().manual_seed(3),
                              dtype=torch.float64)).to(DEV)
    runs = []
    for _ in range(2):
        A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=4)).assembly()
        b = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(1.0, q=4)).assembly()
        A2, b2 = fa.DirichletBC(space, gd=fa.cartesian(u_sin3)).apply(A, b)
        x = fa.cg(A2, b2, rtol=1e-10)
        runs.append((A2.values.clone(), b2.clone(), x.clone()))
    for t0, t1 in zip(*runs):
        assert torch.equal(t0, t1)


@pytest.mark.parametrize("n", [24, 48])
def test_p1_pattern_invariants_large(n):
    """Size-independent properties on bigger Kuhn meshes: nnz = NN + 2 NE,
    symmetry, zero row sums of the constant-coefficient stiffness matrix,
    and pattern == oracle pattern (brute force, n=24)."""
    mesh = TetrahedronMesh.from_box(nx=n, ny=n, nz=n, device=DEV)
    space = LagrangeFESpace(mesh, 1)
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(q=1)).assembly()
    assert A.nnz == mesh.number_of_nodes() + 2 * mesh.number_of_edges()
    ones = torch.ones(A.shape[0], dtype=torch.float64, device=DEV)
    assert float(torch.abs(A @ ones).max()) < 1e-12 * float(torch.abs(A.values).max())
    S = A.to_scipy()
    assert abs(S - S.T).max() < 1e-13 * abs(S).max()
    if n == 24:
        o = tet_from_box(nx=n, ny=n, nz=n)
        c2d = o.cell_to_ipoint(1)
        ref = ofem.assemble_matrix(c2d, [ofem.diffusion_element(o, 1, None, 1)], o.number_of_global_ipoints(1))
        np.testing.assert_array_equal(_np(A.crow), ref.indptr)
        np.testing.assert_array_equal(_np(A.col), ref.indices)
        np.testing.assert_allclose(_np(A.values), ref.data, rtol=0, atol=1e-12 * np.abs(ref.data).max())


def test_fp32_values_match_fp64():
    mesh = TetrahedronMesh.from_box(nx=8, ny=8, nz=8, device=DEV)
    space = LagrangeFESpace(mesh, 1)
    bf = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(2.0, q=2))
    A64 = bf.assembly()
    A32 = bf.assembly(dtype=torch.float32)
    assert A32.values.dtype == torch.float32 and torch.equal(A32.col, A64.col)
    np.testing.assert_allclose(_np(A32.values), _np(A64.values), rtol=0, atol=2e-6 * float(A64.values.abs().max()))
    x = torch.rand(A64.shape[0], dtype=torch.float64, device=DEV)
    np.testing.assert_allclose(_np(A32 @ x.float()), _np(A64 @ x), rtol=0, atol=1e-5 * float((A64 @ x).abs().max()))


@pytest.mark.parametrize("dim", [2, 3])
def test_bc_to_point_matches_host(dim):
    """fa_bc_to_point == the host evaluation sum_v bcs[q, v] node[cell[c, v]]."""
    if dim == 2:
        mesh = fa.TriangleMesh.from_box(nx=7, ny=5, device=DEV)
    else:
        mesh = fa.TetrahedronMesh.from_box(nx=4, ny=3, nz=5, device=DEV)
    bcs, _ = mesh.quadrature_formula(4).get_quadrature_points_and_weights()
    got = mesh.bc_to_point(bcs).cpu().numpy()
    node = mesh.node.cpu().numpy()
    cell = mesh.cell.cpu().numpy()
    ref = np.einsum('qv, cvd -> cqd', np.asarray(bcs.cpu() if torch.is_tensor(bcs) else bcs), node[cell])
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-15)


def test_boundary_data_cached_per_space():
    """Second DirichletBC on the same space reuses the filtered pattern and
    gives identical results."""
    mesh = fa.TetrahedronMesh.from_box(nx=6, ny=6, nz=6, device=DEV)
    space = fa.LagrangeFESpace(mesh, 2)
    u = fa.cartesian(lambda x: x[..., 0] * x[..., 1] + x[..., 2])
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(q=3)).assembly()
    F = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(1.0, q=3)).assembly()
    A1, F1 = fa.DirichletBC(space, gd=u).apply(A, F)
    A2, F2 = fa.DirichletBC(space, gd=u).apply(A, F)
    assert A2.col is A1.col and torch.equal(A1.values, A2.values) and torch.equal(F1, F2)
    ref_flag = space.is_boundary_dof(threshold=lambda p: torch.ones(p.shape[0], dtype=torch.bool, device=p.device),
                                     method='interp')
    assert torch.equal(space.is_boundary_dof(), ref_flag)


@pytest.mark.parametrize("dim,box", [(2, False), (3, False), (3, True)])
def test_fused_p1_matches_element_gather(dim, box, monkeypatch):
    """The fused P1 kernels (element rows computed inside the CSR-row sum),
    and for tetrahedra the structured from_box kernel, against the
    element-tensor + fixed-order gather path, diffusion + mass with a
    per-(cell, point) coefficient and a per-cell source."""
    from fealpy_amd.fem import integrators
    from fealpy_amd.fem.pattern import pattern_of
    monkeypatch.setattr(integrators, "BOX_ASSEMBLY", box)
    if dim == 2:
        mesh = fa.TriangleMesh.from_box(nx=9, ny=7, device=DEV)
    else:
        mesh = fa.TetrahedronMesh.from_box(nx=5, ny=4, nz=6, device=DEV)
    space = fa.LagrangeFESpace(mesh, 1)
    nc = mesh.number_of_cells()
    g = np.random.default_rng(3)
    nq = space.tables(3)[1].shape[0]
    kq = torch.as_tensor(np.exp(g.standard_normal((nc, nq))), device=DEV)
    kc = torch.as_tensor(np.exp(g.standard_normal(nc)), device=DEV)
    di, ma = fa.ScalarDiffusionIntegrator(kq, q=3), fa.ScalarMassIntegrator(kc, q=3)
    A = fa.BilinearForm(space).add_integrator(di, ma).assembly()
    pat = pattern_of(space)
    elem = fa.fem.integrators.GroupIntegrator(di, ma).assembly(space)
    ref = pat.gather_matrix(elem)
    assert torch.equal(A.crow, pat.indptr) and torch.equal(A.col, pat.indices)
    scale = float(ref.abs().max())
    assert float((A.values - ref).abs().max()) <= 1e-13 * scale
    src = fa.ScalarSourceIntegrator(kc, q=3)
    F = fa.LinearForm(space).add_integrator(src).assembly()
    Fref = pat.gather_vector(src.assembly(space))
    assert float((F - Fref).abs().max()) <= 1e-13 * float(Fref.abs().max())
    A32 = fa.BilinearForm(space).add_integrator(di, ma).assembly(dtype=torch.float32)
    assert A32.values.dtype == torch.float32
    assert float((A32.values.double() - ref).abs().max()) <= 1e-6 * scale


@pytest.mark.parametrize("dim", [2, 3])
def test_incidence_two_pass_p1(dim, monkeypatch):
    """Two-pass P1 assembly in incidence order: the maps equal a NumPy
    construction; the FP64 CSR values equal the fused single kernel bit for
    bit (diffusion + mass with per-point / per-cell coefficients); FP32
    values (element rows kept and summed in FP32) within FP32 rounding of
    them; the load vector equals the fused kernel to rounding (that kernel
    adds lane partials by a tree, this one in cell order)."""
    from fealpy_amd import _native as nat
    from fealpy_amd.fem import integrators
    from fealpy_amd.fem.pattern import pattern_of
    monkeypatch.setattr(integrators, "TWO_PASS_MIN_CELLS", 0)   # small mesh, two-pass path
    monkeypatch.setattr(integrators, "BOX_ASSEMBLY", False)     # not the structured from_box kernel
    if dim == 2:
        mesh = fa.TriangleMesh.from_box(nx=7, ny=5, device=DEV)
    else:
        mesh = fa.TetrahedronMesh.from_box(nx=4, ny=3, nz=5, device=DEV)
    space = fa.LagrangeFESpace(mesh, 1)
    pat = pattern_of(space)
    inc, pos = pat.incidence_maps()
    adj = pat.adj.cpu().numpy()
    c2d = pat.cell2dof.cpu().numpy()
    crow, col = pat.indptr.cpu().numpy(), pat.indices.cpu().numpy()
    aptr = pat.adj_ptr.cpu().numpy()
    inc_h, pos_h = inc.cpu().numpy(), pos.cpu().numpy().view(np.uint32)
    ld = pat.ldof
    for r in range(pat.gdof):
        cols = list(col[crow[r]:crow[r + 1]])
        for k in range(aptr[r], aptr[r + 1]):
            c, i = divmod(int(adj[k]), ld)
            assert inc_h[c * ld + i] == k
            for j in range(ld):
                assert (int(pos_h[k]) >> (8 * j)) & 255 == cols.index(c2d[c, j])
    nc = mesh.number_of_cells()
    g = np.random.default_rng(12)
    q = 3
    _, qw, phi, _ = space.tables(q)
    nq = qw.shape[0]
    kq = torch.as_tensor(np.exp(g.standard_normal((nc, nq))), device=DEV)
    kc = torch.as_tensor(np.exp(g.standard_normal(nc)), device=DEV)
    L = nat.lib()
    cell = space.cell_to_dof()
    for dt in (torch.float64, torch.float32):
        two = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(kq, q=q),
                                                    fa.ScalarMassIntegrator(kc, q=q)).assembly(dtype=dt).values
        fused = torch.empty(pat.nnz, dtype=dt, device=DEV)
        nat.check(L.fa_assemble_p1_matrix(nat.dtype_code(dt), dim, pat.gdof, nq, nat.ptr(mesh.node), nat.ptr(cell),
                                          nat.ptr(qw), nat.ptr(phi), nat.COEF_CELLQ, 1.0, nat.ptr(kq), nat.COEF_CELL,
                                          1.0, nat.ptr(kc), nat.ptr(pat.adj_ptr), nat.ptr(pat.adj),
                                          nat.ptr(pat.indptr), nat.ptr(pat.indices), pat.max_row, nat.ptr(fused),
                                          nat.stream()), "fa_assemble_p1_matrix")
        if dt == torch.float64:
            assert torch.equal(two, fused)
            f64 = fused
        else:
            scale = float(f64.abs().max())
            assert float((two.double() - f64).abs().max()) <= 4e-7 * scale
            assert float((fused.double() - f64).abs().max()) <= 1e-7 * scale
    F = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(kc, q=q)).assembly()
    Ff = torch.empty(pat.gdof, dtype=torch.float64, device=DEV)
    nat.check(L.fa_assemble_p1_vector(nat.FA_F64, dim, pat.gdof, nq, nat.ptr(mesh.node), nat.ptr(cell), nat.ptr(qw),
                                      nat.ptr(phi), nat.COEF_CELL, 1.0, nat.ptr(kc), nat.ptr(pat.adj_ptr),
                                      nat.ptr(pat.adj), nat.ptr(Ff), nat.stream()), "fa_assemble_p1_vector")
    assert float((F - Ff).abs().max()) <= 1e-14 * float(Ff.abs().max())


@pytest.mark.parametrize("dim", [2, 3])
def test_dirichlet_apply_one_pass_equals_two(dim):
    """DirichletBC.apply (one pass over A) equals apply_vector + apply_matrix
    bit for bit, FP64 and FP32, CSR and COO input."""
    if dim == 2:
        mesh = fa.TriangleMesh.from_box(nx=9, ny=6, device=DEV)
        u = fa.cartesian(lambda p: torch.sin(np.pi * p[..., 0]) * p[..., 1] + 1.0)
    else:
        mesh = fa.TetrahedronMesh.from_box(nx=4, ny=5, nz=3, device=DEV)
        u = fa.cartesian(lambda p: torch.cos(p[..., 0]) + p[..., 1] * p[..., 2])
    space = fa.LagrangeFESpace(mesh, 2)
    k = torch.as_tensor(np.exp(np.random.default_rng(4).standard_normal(mesh.number_of_cells())), device=DEV)
    for dt in (torch.float64, torch.float32):
        A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=4)).assembly(dtype=dt)
        F = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(1.0, q=4)).assembly(dtype=dt)
        bc = fa.DirichletBC(space, gd=u)
        A1, F1 = bc.apply(A, F)
        F2 = bc.apply_vector(F, A)
        A2 = bc.apply_matrix(A)
        assert torch.equal(A1.crow, A2.crow) and torch.equal(A1.col, A2.col)
        assert torch.equal(A1.values, A2.values) and torch.equal(F1, F2), dt
        Ac, Fc = bc.apply(A.tocoo(), F)
        assert torch.equal(Ac.tocsr().values, A2.values) and torch.equal(Fc, F2)
