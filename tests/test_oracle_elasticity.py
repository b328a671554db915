"""Oracle for vector P2 linear elasticity (CPU).

Pinned by the reference-written fixture (tests/golden/fixtures_elasticity3d_p2.npz:
element matrices, assembled CSR, vector load) and by what the mathematics
fixes: rigid motions span the kernel of every element matrix, the element
matrix is symmetric, and with lam = 0, mu = 1 its diagonal blocks are
A_aa + sum_b A_bb -- a transposed block or a wrong layout fails one of these."""
import numpy as np
import pytest

from oracle import fem as ofem
from oracle.mesh import tet_from_box, tri_from_box

PI = np.pi


def _body(lam, mu):
    c = np.array([1.0, 2.0, 3.0])

    def f(p):
        x, y, z = p[..., 0], p[..., 1], p[..., 2]
        s = np.sin(PI * x) * np.sin(PI * y) * np.sin(PI * z)
        H = np.zeros(p.shape[:-1] + (3, 3))
        H[..., 0, 0] = H[..., 1, 1] = H[..., 2, 2] = -PI ** 2 * s
        H[..., 0, 1] = H[..., 1, 0] = PI ** 2 * np.cos(PI * x) * np.cos(PI * y) * np.sin(PI * z)
        H[..., 0, 2] = H[..., 2, 0] = PI ** 2 * np.cos(PI * x) * np.sin(PI * y) * np.cos(PI * z)
        H[..., 1, 2] = H[..., 2, 1] = PI ** 2 * np.sin(PI * x) * np.cos(PI * y) * np.cos(PI * z)
        return 3 * PI ** 2 * mu * s[..., None] * c - (lam + mu) * np.einsum('...ij, j -> ...i', H, c)
    return ofem.cartesian(f)


def test_elasticity_matches_reference_fixture(golden):
    g = golden("fixtures_elasticity3d_p2.npz")
    n, lam, mu = int(g["n"]), float(g["lam"]), float(g["mu"])
    mesh = tet_from_box(nx=n, ny=n, nz=n)
    c2d = mesh.cell_to_ipoint(2)
    sg = mesh.number_of_global_ipoints(2)
    t2d = ofem.tensor_cell2dof(c2d, 3, sg, False)
    np.testing.assert_array_equal(t2d, g["cell2dof"])
    assert 3 * sg == int(g["gdof"])
    ke = ofem.elasticity_element(mesh, 2, lam, mu, q=5)
    np.testing.assert_allclose(ke, g["ke"], rtol=0, atol=1e-12 * np.abs(g["ke"]).max())
    A = ofem.assemble_matrix(t2d, [ke], 3 * sg).canonical()
    np.testing.assert_array_equal(A.indptr, g["A_indptr"])
    np.testing.assert_array_equal(A.indices, g["A_indices"])
    np.testing.assert_allclose(A.data, g["A_data"], rtol=0, atol=1e-12 * np.abs(g["A_data"]).max())
    fe = ofem.vector_source_element(mesh, 2, _body(lam, mu), q=5)
    F = ofem.assemble_vector(t2d, [fe], 3 * sg)
    np.testing.assert_allclose(F, g["F"], rtol=0, atol=1e-12 * np.abs(g["F"]).max())


@pytest.mark.parametrize("dim,p,prio", [(3, 1, False), (3, 2, False), (3, 2, True), (3, 4, True), (2, 2, False), (2, 1, True),
                                         (2, 4, False)])
def test_rigid_motions_are_in_the_kernel(dim, p, prio):
    mesh = tet_from_box(nx=2, ny=1, nz=2) if dim == 3 else tri_from_box(nx=3, ny=2)
    ke = ofem.elasticity_element(mesh, p, 1.3, 0.7, q=p + 2, dof_priority=prio)
    ip = mesh.interpolation_points(p)[mesh.cell_to_ipoint(p)]            # (NC, ldof, GD)
    NC, ldof, GD = ip.shape
    modes = [np.broadcast_to(np.eye(GD)[a], ip.shape) for a in range(GD)]
    if GD == 3:
        for w in np.eye(3):
            modes.append(np.cross(w, ip))
    else:
        modes.append(np.stack([-ip[..., 1], ip[..., 0]], axis=-1))
    for u in modes:
        v = u.swapaxes(1, 2).reshape(NC, -1) if prio else u.reshape(NC, -1)
        r = np.einsum('cij, cj -> ci', ke, v)
        assert np.abs(r).max() < 1e-12 * np.abs(ke).max()
    np.testing.assert_allclose(ke, ke.swapaxes(1, 2), rtol=0, atol=1e-14 * np.abs(ke).max())


def test_shear_only_blocks():
    """lam = 0, mu = 1: diagonal block (a, a) = A_aa + sum_b A_bb, off-diagonal (a, b) = A_ba."""
    mesh = tet_from_box(nx=1, ny=1, nz=1)
    ke = ofem.elasticity_element(mesh, 1, 0.0, 1.0, q=2)
    bcs, ws = mesh.quadrature_formula(2)
    g = mesh.grad_shape_function(bcs, 1, 'x')
    cm = mesh.entity_measure('cell')
    A = np.einsum('q, cqia, cqjb, c -> cabij', ws, g, g, cm)
    tr = A[:, 0, 0] + A[:, 1, 1] + A[:, 2, 2]
    for a in range(3):
        np.testing.assert_allclose(ke[:, a::3, a::3], A[:, a, a] + tr, atol=1e-15)
        for b in range(3):
            if a != b:
                np.testing.assert_allclose(ke[:, a::3, b::3], A[:, b, a], atol=1e-15)


@pytest.mark.parametrize("dim,p,prio", [(3, 1, False), (3, 3, True), (3, 4, False), (2, 1, True), (2, 3, False),
                                         (2, 4, True)])
def test_energy_of_linear_displacement(dim, p, prio):
    """u = G x is represented exactly for every p, so u^T K u summed over the
    cells is |Omega| (2 mu eps:eps + lam tr(eps)^2) with eps = sym(G)
    (3D and plane strain): pins the block coefficients and |T| for any p."""
    mesh = tet_from_box(nx=2, ny=1, nz=1) if dim == 3 else tri_from_box(nx=3, ny=2)
    lam, mu = 1.3, 0.7
    ke = ofem.elasticity_element(mesh, p, lam, mu, q=2 * p, dof_priority=prio)
    ip = mesh.interpolation_points(p)[mesh.cell_to_ipoint(p)]            # (NC, ldof, GD)
    NC, ldof, GD = ip.shape
    G = np.random.default_rng(7).standard_normal((GD, GD))
    u = ip @ G.T
    v = u.swapaxes(1, 2).reshape(NC, -1) if prio else u.reshape(NC, -1)
    energy = np.einsum('ci, cij, cj ->', v, ke, v)
    eps = 0.5 * (G + G.T)
    exact = 1.0 * (2 * mu * np.sum(eps * eps) + lam * np.trace(eps) ** 2)   # |Omega| = 1
    assert abs(energy - exact) <= 1e-12 * abs(exact)
