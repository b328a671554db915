"""Pin the oracle's integrators, COO->CSR assembly, Dirichlet BC, CG and L2
error against fixtures written by running the reference (tests/golden/
make_fixtures.py) and against closed forms."""
import numpy as np
import pytest

from oracle import fem
from oracle.mesh import tet_from_box, tri_from_box

PI = np.pi


def _csr(F, prefix):
    return fem.CSR(F[prefix + "_indptr"], F[prefix + "_indices"], F[prefix + "_data"],
                   (len(F[prefix + "_indptr"]) - 1,) * 2)


def _assert_csr(A, F, prefix, rtol=1e-12):
    A = A.canonical()
    np.testing.assert_array_equal(A.indptr, F[prefix + "_indptr"])
    np.testing.assert_array_equal(A.indices, F[prefix + "_indices"])
    scale = np.abs(F[prefix + "_data"]).max()
    np.testing.assert_allclose(A.data, F[prefix + "_data"], rtol=0, atol=rtol * scale)


def _solve_and_check(mesh, p, A, Fv, u, F, q_err=3):
    c2d = mesh.cell_to_ipoint(p)
    is_bd = mesh.is_boundary_dof(p)
    np.testing.assert_array_equal(is_bd, F["is_bd"])
    Abc, Fbc = fem.dirichlet_apply(A, Fv, is_bd, mesh.interpolation_points(p), u)
    _assert_csr(Abc, F, "Abc")
    np.testing.assert_allclose(Fbc, F["Fbc"], rtol=0, atol=1e-12 * np.abs(F["Fbc"]).max())
    x, info = fem.cg(Abc, Fbc, atol=1e-14, rtol=1e-14, maxit=20000)
    np.testing.assert_allclose(x, F["x"], rtol=0, atol=1e-10 * np.abs(F["x"]).max())
    err = fem.l2_error(mesh, p, u, x, c2d, q=q_err)
    assert abs(err - float(F["l2_error"])) <= 1e-6 * float(F["l2_error"])
    return Abc, Fbc, x, info


def test_poisson2d_p1_full(golden):
    """Tri 64x64 P1, q=2: A, F, BC'd system, CG (rtol 1e-10 iteration count), L2 error."""
    F = golden("fixtures_poisson2d_p1.npz")
    mesh = tri_from_box((0, 1, 0, 1), 64, 64)
    c2d = mesh.cell_to_ipoint(1)
    np.testing.assert_array_equal(c2d, F["cell2dof"])
    u = fem.cartesian(lambda p: np.sin(PI * p[..., 0]) * np.sin(PI * p[..., 1]))
    f = fem.cartesian(lambda p: 2 * PI ** 2 * np.sin(PI * p[..., 0]) * np.sin(PI * p[..., 1]))
    A = fem.assemble_matrix(c2d, [fem.diffusion_element(mesh, 1, None, 2)], mesh.number_of_global_ipoints(1))
    _assert_csr(A, F, "A")
    Fv = fem.assemble_vector(c2d, [fem.source_element(mesh, 1, f, 2)], mesh.number_of_global_ipoints(1))
    np.testing.assert_allclose(Fv, F["F"], rtol=0, atol=1e-13 * np.abs(F["F"]).max())
    Abc, Fbc, _, _ = _solve_and_check(mesh, 1, A, Fv, u, F)
    x2, info2 = fem.cg(Abc, Fbc, rtol=1e-10)
    assert info2["niter"] == int(F["cg_niter"])
    np.testing.assert_allclose(x2, F["x_cg"], rtol=0, atol=1e-12)


def test_diffusion3d_p1_lognormal(golden):
    F = golden("fixtures_diffusion3d_p1_lognormal.npz")
    n, k = int(F["n"]), F["k"]
    mesh = tet_from_box(nx=n, ny=n, nz=n)
    c2d = mesh.cell_to_ipoint(1)
    u = fem.cartesian(lambda p: np.sin(PI * p[..., 0]) * np.sin(PI * p[..., 1]) * np.sin(PI * p[..., 2]))
    f = fem.barycentric(lambda bcs, index=None: 3 * PI ** 2 * k[:, None] * u(mesh.bc_to_point(bcs)))
    gdof = mesh.number_of_global_ipoints(1)
    A = fem.assemble_matrix(c2d, [fem.diffusion_element(mesh, 1, k, 2)], gdof)
    _assert_csr(A, F, "A")
    Fv = fem.assemble_vector(c2d, [fem.source_element(mesh, 1, f, 2)], gdof)
    np.testing.assert_allclose(Fv, F["F"], rtol=0, atol=1e-13 * np.abs(F["F"]).max())
    _solve_and_check(mesh, 1, A, Fv, u, F)


def _k_smooth(p):
    return 1 + 0.5 * np.sin(2 * PI * p[..., 0]) * np.cos(2 * PI * p[..., 1]) * p[..., 2]


def _f_highorder(pt):
    x, y, z = pt[..., 0], pt[..., 1], pt[..., 2]
    sx, sy, sz = np.sin(PI * x), np.sin(PI * y), np.sin(PI * z)
    cx, cy, cz = np.cos(PI * x), np.cos(PI * y), np.cos(PI * z)
    uu = sx * sy * sz
    gk = (PI * np.cos(2 * PI * x) * np.cos(2 * PI * y) * z, -PI * np.sin(2 * PI * x) * np.sin(2 * PI * y) * z,
          0.5 * np.sin(2 * PI * x) * np.cos(2 * PI * y))
    gu = (PI * cx * sy * sz, PI * sx * cy * sz, PI * sx * sy * cz)
    return 3 * PI ** 2 * _k_smooth(pt) * uu - sum(a * b for a, b in zip(gk, gu)) + uu


@pytest.mark.parametrize("p", [3, 4])
def test_highorder3d(golden, p):
    F = golden(f"fixtures_highorder3d_p{p}.npz")
    n, q = int(F["n"]), int(F["q"])
    mesh = tet_from_box(nx=n, ny=n, nz=n)
    c2d = mesh.cell_to_ipoint(p)
    np.testing.assert_array_equal(c2d, F["cell2dof"])
    gdof = mesh.number_of_global_ipoints(p)
    K = fem.diffusion_element(mesh, p, fem.cartesian(_k_smooth), q) + fem.mass_element(mesh, p, 1.0, q)
    A = fem.assemble_matrix(c2d, [K], gdof)
    _assert_csr(A, F, "A")
    Fv = fem.assemble_vector(c2d, [fem.source_element(mesh, p, fem.cartesian(_f_highorder), q)], gdof)
    np.testing.assert_allclose(Fv, F["F"], rtol=0, atol=1e-13 * np.abs(F["F"]).max())
    u = fem.cartesian(lambda pt: np.sin(PI * pt[..., 0]) * np.sin(PI * pt[..., 1]) * np.sin(PI * pt[..., 2]))
    _solve_and_check(mesh, p, A, Fv, u, F, q_err=q)


def test_poisson2d_p2_fourier(golden):
    F = golden("fixtures_poisson2d_p2_fourier.npz")
    n, k, a, km, lm = int(F["n"]), F["k"], F["a"], F["km"], F["lm"]
    mesh = tri_from_box((0, 1, 0, 1), n, n)
    c2d = mesh.cell_to_ipoint(2)
    np.testing.assert_array_equal(c2d, F["cell2dof"])
    f = fem.cartesian(lambda p: sum(a[m] * np.sin(km[m] * PI * p[..., 0]) * np.sin(lm[m] * PI * p[..., 1])
                                    for m in range(8)))
    gdof = mesh.number_of_global_ipoints(2)
    A = fem.assemble_matrix(c2d, [fem.diffusion_element(mesh, 2, k)], gdof)
    _assert_csr(A, F, "A")
    Fv = fem.assemble_vector(c2d, [fem.source_element(mesh, 2, f)], gdof)
    np.testing.assert_allclose(Fv, F["F"], rtol=0, atol=1e-13 * np.abs(F["F"]).max())
    zero = fem.cartesian(lambda p: np.zeros(p.shape[:-1]))
    _solve_and_check(mesh, 2, A, Fv, zero, F)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_tet_patterns(golden, p):
    """Pattern of the tet 4^3 P1..P4 forms: the reference's, and the closed-form
    gdof/nnz 125/1333, 729/17025, 2197/91429, 4913/327073."""
    F = golden("fixtures_patterns.npz")
    mesh = tet_from_box(nx=4, ny=4, nz=4)
    c2d = mesh.cell_to_ipoint(p)
    np.testing.assert_array_equal(c2d, F[f"tet4_p{p}_cell2dof"])
    gdof = mesh.number_of_global_ipoints(p)
    assert (gdof, int(F[f"tet4_p{p}_nnz"])) == {1: (125, 1333), 2: (729, 17025),
                                                  3: (2197, 91429), 4: (4913, 327073)}[p]
    A = fem.assemble_matrix(c2d, [fem.diffusion_element(mesh, p, None, p + 1)], gdof)
    _assert_csr(A, F, f"tet4_p{p}")
    # brute force pattern: set of (i, j) over cells
    pairs = {(int(i), int(j)) for row in c2d for i in row for j in row}
    assert len(pairs) == A.indptr[-1]
    # row sums of a stiffness matrix vanish (constants are in the kernel)
    assert np.abs(A.todense().sum(1)).max() < 1e-12 * np.abs(A.data).max()


def test_p1_closed_forms():
    """P1 element matrices: K_e = |T| k grad(l_i).grad(l_j) * sum(w); mass
    |T|(1 + delta_ij)/((d+1)(d+2)) for an exact rule."""
    mesh = tet_from_box(nx=2, ny=1, nz=1)
    g = mesh.grad_lambda()
    vol = mesh.entity_measure('cell')
    for q in (1, 2, 3):
        _, ws = mesh.quadrature_formula(q)
        K = fem.diffusion_element(mesh, 1, 2.5, q)
        np.testing.assert_allclose(K, 2.5 * ws.sum() * vol[:, None, None] * np.einsum('cim, cjm -> cij', g, g),
                                   rtol=1e-13, atol=1e-15)
    M = fem.mass_element(mesh, 1, None, 2)
    want = vol[:, None, None] * (np.ones((4, 4)) + np.eye(4)) / 20
    np.testing.assert_allclose(M, want, rtol=1e-13)


def test_cg_matches_direct_solve():
    """Textbook CG on a small SPD system reaches scipy's direct solution."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    mesh = tri_from_box((0, 1, 0, 1), 6, 6)
    c2d = mesh.cell_to_ipoint(1)
    A = fem.assemble_matrix(c2d, [fem.diffusion_element(mesh, 1, None, 2) + fem.mass_element(mesh, 1, 1.0, 2)],
                            mesh.number_of_global_ipoints(1))
    b = np.random.default_rng(1).standard_normal(A.shape[0])
    x, info = fem.cg(A, b, atol=1e-15, rtol=1e-15, maxit=1000)
    S = sp.csr_matrix((A.data, A.indices, A.indptr), shape=A.shape)
    np.testing.assert_allclose(x, spla.spsolve(S.tocsc(), b), rtol=0, atol=1e-11)


def _spd_matrix(n=6):
    mesh = tri_from_box((0, 1, 0, 1), n, n)
    c2d = mesh.cell_to_ipoint(1)
    return fem.assemble_matrix(c2d, [fem.diffusion_element(mesh, 1, None, 2) + fem.mass_element(mesh, 1, 1.0, 2)],
                               mesh.number_of_global_ipoints(1))


def test_cg_batch_columns_reach_direct_solves():
    """Right-hand sides as columns (cg.py:85-113, lockstep, alpha and beta per
    column): every column reaches scipy's direct solution."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    A = _spd_matrix()
    B = np.random.default_rng(2).standard_normal((A.shape[0], 3)) * np.array([1.0, 10.0, 0.1])
    X, info = fem.cg(A, B, atol=1e-15, rtol=1e-15, maxit=1000)
    S = sp.csr_matrix((A.data, A.indices, A.indptr), shape=A.shape).tocsc()
    for k in range(3):
        np.testing.assert_allclose(X[:, k], spla.spsolve(S, B[:, k]), rtol=0, atol=1e-11 * np.abs(B[:, k]).max())


def test_cg_batch_of_identical_columns_equals_single_solve():
    """Identical columns have the single solve's alpha and beta, and the
    relative test sqrt(sum_k r.z) < rtol ||B||_F is the single test scaled by
    sqrt(B) on both sides: same iterates (up to the summation order of the
    column dots), same iteration count."""
    A = _spd_matrix()
    b = np.random.default_rng(3).standard_normal(A.shape[0])
    x1, i1 = fem.cg(A, b, atol=0.0, rtol=1e-6)
    X, info = fem.cg(A, np.stack([b] * 3, axis=1), atol=0.0, rtol=1e-6)
    assert info['niter'] == i1['niter']
    for k in range(3):
        np.testing.assert_allclose(X[:, k], x1, rtol=0, atol=1e-8 * np.abs(x1).max())


def test_cg_batch_absolute_test_sums_the_columns():
    """The absolute test sums r.z over the columns: 64 identical columns stop
    where a single solve with atol / 8 stops (a maximum or mean over the
    columns would stop where the single solve with atol stops)."""
    A = _spd_matrix()
    b = np.random.default_rng(4).standard_normal(A.shape[0])
    a = 1e-6 * np.linalg.norm(b)
    _, i_eighth = fem.cg(A, b, atol=a / 8, rtol=0.0)
    _, i_full = fem.cg(A, b, atol=a, rtol=0.0)
    X, info = fem.cg(A, np.stack([b] * 64, axis=1), atol=a, rtol=0.0)
    assert i_full['niter'] < i_eighth['niter']
    assert info['niter'] == i_eighth['niter']


def test_l2_error_converges_second_order():
    u = fem.cartesian(lambda p: np.sin(PI * p[..., 0]) * np.sin(PI * p[..., 1]))
    f = fem.cartesian(lambda p: 2 * PI ** 2 * np.sin(PI * p[..., 0]) * np.sin(PI * p[..., 1]))
    errs = []
    for n in (8, 16, 32):
        mesh = tri_from_box((0, 1, 0, 1), n, n)
        c2d = mesh.cell_to_ipoint(1)
        gdof = mesh.number_of_global_ipoints(1)
        A = fem.assemble_matrix(c2d, [fem.diffusion_element(mesh, 1, None, 3)], gdof)
        Fv = fem.assemble_vector(c2d, [fem.source_element(mesh, 1, f, 3)], gdof)
        Abc, Fbc = fem.dirichlet_apply(A, Fv, mesh.is_boundary_dof(1), mesh.interpolation_points(1), u)
        x, _ = fem.cg(Abc, Fbc, atol=1e-14, rtol=1e-14)
        errs.append(fem.l2_error(mesh, 1, u, x, c2d))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > 1.9), rates
