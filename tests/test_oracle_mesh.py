"""Pin the oracle's mesh / basis / quadrature layer against the reference's own
golden vectors (parsed out of test/*/…_data.py) and against closed forms."""
import numpy as np
import pytest

from oracle.basis import multi_index_matrix, quadrature, simplex_grad_shape_function, \
    simplex_shape_function, stroud
from oracle.mesh import tet_from_box, tri_from_box


def test_tet_from_box_threshold(golden):
    """test/mesh/test_tetrahedron_mesh.py:66-87 (+ data :4618-4947)."""
    T = golden("ref_tetrahedron_mesh.npz")
    th = lambda p: (p[..., 0] < 0.5) & (p[..., 1] < 0.5) & (p[..., 2] < 0.5)
    m = tet_from_box(nx=3, ny=2, nz=1, threshold=th)
    np.testing.assert_allclose(m.node, T["from_box::0::node"], atol=1e-14)
    for k in ("edge", "face", "cell", "face2cell"):
        np.testing.assert_array_equal(getattr(m, k), T["from_box::0::" + k])
    assert (m.number_of_nodes(), m.number_of_edges(), m.number_of_faces(), m.number_of_cells()) == \
        tuple(int(T[f"from_box::0::{k}"]) for k in ("NN", "NE", "NF", "NC"))


def test_tet_measures_and_grad_lambda(golden):
    """test_tetrahedron_mesh.py:89-110."""
    T = golden("ref_tetrahedron_mesh.npz")
    m = tet_from_box(nx=3, ny=2, nz=1)
    np.testing.assert_allclose(m.entity_measure('edge'), T["entity_measure::0::em"], atol=1e-14)
    np.testing.assert_allclose(m.entity_measure('face'), T["entity_measure::0::fm"], atol=1e-14)
    np.testing.assert_allclose(m.entity_measure('cell'), T["entity_measure::0::cm"], atol=1e-14)
    np.testing.assert_allclose(m.grad_lambda(), T["grad_lambda::0::glambda"], atol=1e-14)


def test_tet_grad_shape_function(golden):
    """test_tetrahedron_mesh.py:112-121.  The stored 'x' array is laid out
    (NQ, NC, ldof, GD) -- an older layout; the current reference returns
    (NC, NQ, ldof, GD) (mesh_base.py:744), so it is compared transposed."""
    T = golden("ref_tetrahedron_mesh.npz")
    m = tet_from_box(nx=1, ny=1, nz=1)
    bcs, _ = quadrature(3, 3)
    np.testing.assert_allclose(m.grad_shape_function(bcs, 2, 'x').transpose(1, 0, 2, 3),
                               T["grad_shape_function::0::grad_shape_function"], atol=1e-14)
    np.testing.assert_allclose(m.grad_shape_function(bcs, 2, 'u'),
                               T["grad_shape_function::1::grad_shape_function"], atol=1e-14)


def test_tet_ipoints_and_dof_maps(golden):
    """test_tetrahedron_mesh.py:123-163: gdof(p=4)=585, ipoints, face/cell->ipoint (p=4)."""
    T = golden("ref_tetrahedron_mesh.npz")
    m = tet_from_box(nx=3, ny=2, nz=1)
    assert m.number_of_global_ipoints(4) == int(T["number_of_global_ipoints::0::gdof"])
    for i, want in ((0, 20), (1, 10), (2, 2)):
        assert int(T[f"number_of_local_ipoints::{i}::ldof"]) == want
    np.testing.assert_allclose(m.interpolation_points(4), T["interpolation_points::0::ipoint"], atol=1e-14)
    np.testing.assert_array_equal(m.face_to_ipoint(4), T["face_to_ipoint::0::f2p"])
    np.testing.assert_array_equal(m.cell_to_ipoint(4), T["cell_to_ipoint::0::cell2ipoint"])


def test_tri_topology_and_ipoints(golden):
    """test/mesh/test_triangle_mesh.py:57-126."""
    R = golden("ref_triangle_mesh.npz")
    m = tri_from_box(nx=2, ny=2)
    for k in ("cell", "edge", "face2cell"):
        np.testing.assert_array_equal(getattr(m, k), R["from_box_data::0::" + k])
    np.testing.assert_allclose(m.node, R["from_box_data::0::node"], atol=1e-14)
    m2 = tri_from_box((-1, 1, -1, 1), 2, 2)
    np.testing.assert_allclose(m2.grad_lambda(), R["grad_lambda_data::0::val"], atol=1e-14)
    np.testing.assert_allclose(m.interpolation_points(4), R["interpolation_point_data::0::ips"], atol=1e-14)
    np.testing.assert_array_equal(m.cell_to_ipoint(4), R["interpolation_point_data::0::cip"])
    np.testing.assert_array_equal(m.face_to_ipoint(4), R["interpolation_point_data::0::fip"])


def test_tri_grad_shape_function(golden):
    """test_triangle_mesh.py:102-111.  The stored values were produced with a
    more precise 6-point rule than the 16-digit table the reference now ships
    (differences ~5e-9), so the tolerance here is 1e-8, not 1e-14."""
    R = golden("ref_triangle_mesh.npz")
    m = tri_from_box(nx=2, ny=2)
    bcs, _ = quadrature(2, 3)
    np.testing.assert_allclose(m.grad_shape_function(bcs, 2), R["grad_shape_function_data::0::gphi"], atol=1e-8)


def test_lagrange_space_p2(golden):
    """test/functionspace/test_lagrange_fe_space.py:14-59."""
    L = golden("ref_lagrange_fe_space.npz")
    m = tri_from_box((0, 1, 0, 1), 1, 1)
    np.testing.assert_array_equal(m.cell_to_ipoint(2), L["triangle_mesh_one_box::0::cell_to_dof"])
    np.testing.assert_array_equal(m.face_to_ipoint(2), L["triangle_mesh_one_box::0::face_to_dof"])
    np.testing.assert_array_equal(m.is_boundary_dof(2), L["triangle_mesh_one_box::0::is_boundary_dof"])
    bcs = np.array([[1 / 3, 1 / 3, 1 / 3], [0.5, 0.5, 0.0], [0.0, 0.0, 1.0]])
    bcs = L["triangle_mesh_one_box::0::basis"].shape and bcs
    assert m.number_of_global_ipoints(2) == int(L["triangle_mesh_one_box::0::number_of_global_dofs"])


@pytest.mark.parametrize("i", range(8))
def test_stroud(golden, i):
    """test/quadrature/test_stroud_quadrature.py:9-19."""
    S = golden("ref_stroud_quadrature.npz")
    bcs, ws = stroud(int(S[f"data::{i}::dim"]), int(S[f"data::{i}::p"]))
    np.testing.assert_allclose(bcs, S[f"data::{i}::bcs"], atol=1e-15)
    np.testing.assert_allclose(ws, S[f"data::{i}::ws"], atol=1e-15)


@pytest.mark.parametrize("TD,q", [(2, q) for q in range(1, 12)] + [(3, q) for q in range(1, 10)])
def test_quadrature_exactness(TD, q):
    """Closed form: int_T lambda^alpha / |T| = alpha! TD! / (|alpha| + TD)!  for the
    degree the rule claims (tri index->degree per triangle.py comments, tet per
    tetrahedron.py comments, Stroud(n) exact to 2n-1)."""
    from math import factorial
    tri_deg = {1: 1, 2: 2, 3: 4, 4: 5, 5: 7, 6: 8, 7: 10, 8: 12, 9: 14, 10: 2 * 10 - 1, 11: 2 * 11 - 1}
    tet_deg = {1: 1, 2: 2, 3: 3, 4: 5, 5: 6, 6: 8, 7: 9, 8: 15, 9: 17}
    deg = (tri_deg if TD == 2 else tet_deg)[q]
    bcs, ws = quadrature(TD, q)
    assert abs(ws.sum() - 1) < 5e-14 and np.allclose(bcs.sum(1), 1, atol=1e-14)
    for a in multi_index_matrix(deg, TD):
        exact = np.prod([factorial(int(k)) for k in a]) * factorial(TD) / factorial(int(a.sum()) + TD)
        got = np.sum(ws * np.prod(bcs ** a, axis=1))
        assert abs(got - exact) < 5e-13, (a, got, exact)


def test_product_rule_tables_equal_the_oracle_copy():
    """The product's typed-in rule tables against the oracle's copy, which was
    parsed from fealpy/quadrature/{triangle,tetrahedron}.py: same points, same
    order, same rounding."""
    from fealpy_amd.quadrature.tables import TETRAHEDRON_RULES, TRIANGLE_RULES
    from oracle.basis import _TET, _TRI
    assert sorted(TRIANGLE_RULES) == sorted(_TRI) and sorted(TETRAHEDRON_RULES) == sorted(_TET)
    for mine, ref in ((TRIANGLE_RULES, _TRI), (TETRAHEDRON_RULES, _TET)):
        for q in ref:
            np.testing.assert_array_equal(np.array(mine[q], dtype=np.float64), ref[q])


@pytest.mark.parametrize("TD,p", [(2, p) for p in range(1, 5)] + [(3, p) for p in range(1, 5)])
def test_shape_function_kronecker_and_gradient(TD, p):
    """phi_i(x_j) = delta_ij at the lattice points m/p; sum_i phi_i = 1; the
    lambda-gradient matches a central difference of phi."""
    mi = multi_index_matrix(p, TD)
    pts = mi / p
    phi = simplex_shape_function(pts.astype(float), p, mi)
    np.testing.assert_allclose(phi, np.eye(len(mi)), atol=1e-13)
    rng = np.random.default_rng(0)
    bc = rng.dirichlet(np.ones(TD + 1), size=5)
    np.testing.assert_allclose(simplex_shape_function(bc, p, mi).sum(-1), 1, atol=1e-13)
    R = simplex_grad_shape_function(bc, p, mi)
    h = 1e-6
    for m in range(TD + 1):
        e = np.zeros(TD + 1); e[m] = h
        fd = (simplex_shape_function(bc + e, p, mi) - simplex_shape_function(bc - e, p, mi)) / (2 * h)
        np.testing.assert_allclose(R[..., m], fd, atol=1e-6)


@pytest.mark.parametrize("n", [1, 2, 3])
def test_kuhn_counts_closed_form(n):
    """Euler characteristic and the Kuhn counts NE = 7n^3+9n^2+3n... via V-E+F-C = 1."""
    m = tet_from_box(nx=n, ny=n, nz=n)
    V, E, F, C = m.number_of_nodes(), m.number_of_edges(), m.number_of_faces(), m.number_of_cells()
    assert V == (n + 1) ** 3 and C == 6 * n ** 3
    assert V - E + F - C == 1
    assert int(m.boundary_face_flag().sum()) == 12 * n * n
