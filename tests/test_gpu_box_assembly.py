"""The structured P1 kernels of TetrahedronMesh.from_box meshes
(fa_p1_box_matrix, fa_p1_box_vector), against the two-pass incidence
kernels, which
the reference's fixtures pin (the 24^3 lognormal and 128^3 summary tests
run both paths against the reference):

- FP64 values equal to the two-pass kernels' to a few units in the last
  place (every entry is the sum of its cell contributions in ascending cell
  order in both; the element geometry of an axis-aligned cube is the Kuhn
  closed form, elsewhere the same cross products, which the compiler may
  contract differently) and bitwise run to run: one-cube-thick and uneven
  grids, box extents other than the unit cube, diffusion + mass with
  per-point and per-cell coefficients, and a separable expression evaluated
  in the kernel;
- FP32 values are the FP64 sums rounded once;
- load vectors bit for bit equal to the two-pass kernels for per-cell,
  per-point and constant sources; separable sources, taken from per-grid-line
  tables, to a few units in the last place, also on a grid whose nodes were
  moved off the grid lines (those cells evaluate the source directly);
- the CSR pattern built from the grid equals the incidence-list pattern;
- a CSR array that is not the grid's P1 pattern is refused before any launch;
- meshes that are not a full from_box grid (threshold) keep the other kernels."""
import numpy as np
import pytest
import torch

import fealpy_amd as fa
from fealpy_amd import _native as nat
from fealpy_amd.expr import cos, exp, sin, x, y, z
from fealpy_amd.fem import integrators
from fealpy_amd.fem.pattern import pattern_of
from fealpy_amd.mesh import TetrahedronMesh

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def _assemble(space, its, dtype=torch.float64):
    return fa.BilinearForm(space).add_integrator(*its).assembly(dtype=dtype).values


@pytest.mark.parametrize("dims,box", [((1, 1, 1), (0, 1, 0, 1, 0, 1)), ((4, 3, 5), (0, 1, 0, 1, 0, 1)),
                                      ((1, 6, 2), (-3, 5, 0, 2, -1, 1)), ((9, 7, 8), (0, 2, 0, 1, 0, 3))])
def test_box_equals_two_pass(dims, box, monkeypatch):
    mesh = TetrahedronMesh.from_box(list(box), nx=dims[0], ny=dims[1], nz=dims[2], device=DEV)
    space = fa.LagrangeFESpace(mesh, 1)
    pat = pattern_of(space)
    assert integrators._box_grid(space, pat) == dims
    nc = mesh.number_of_cells()
    g = np.random.default_rng(sum(dims))
    q = 3
    nq = space.tables(q)[1].shape[0]
    kq = torch.as_tensor(np.exp(g.standard_normal((nc, nq))), device=DEV)
    kc = torch.as_tensor(np.exp(g.standard_normal(nc)), device=DEV)
    kx = 1.5 + sin(3 * x) * exp(0.1 * y) * z ** 2 + cos(2 * z)
    cases = [[fa.ScalarDiffusionIntegrator(kc, q=2)],
             [fa.ScalarDiffusionIntegrator(kq, q=q), fa.ScalarMassIntegrator(kc, q=q)],
             [fa.ScalarDiffusionIntegrator(kx, q=q), fa.ScalarMassIntegrator(0.25, q=q)],
             [fa.ScalarMassIntegrator(kq, q=q)]]
    for its in cases:
        box64 = _assemble(space, its)
        box32 = _assemble(space, its, torch.float32)
        monkeypatch.setattr(integrators, "BOX_ASSEMBLY", False)
        monkeypatch.setattr(integrators, "TWO_PASS_MIN_CELLS", 0)
        two = _assemble(space, its)
        monkeypatch.setattr(integrators, "BOX_ASSEMBLY", True)
        # the same element rows and the same summation order; the compiler may
        # contract multiply-adds inside the element geometry differently in
        # the two kernels, so entries agree to a few units in the last place
        assert float((box64 - two).abs().max()) <= 1e-15 * float(two.abs().max())
        assert torch.equal(box32, box64.to(torch.float32))
        again = _assemble(space, its)
        assert torch.equal(again, box64)                      # run to run: bitwise


@pytest.mark.parametrize("moved", ["none", "one", "many"])
def test_box_matrix_off_the_grid_lines(moved, monkeypatch):
    """Cubes that are axis-aligned boxes take the closed-form Kuhn geometry,
    cubes with a moved corner the cross-product geometry; either way the
    two-pass kernels' values to a few units in the last place."""
    dims = (5, 4, 6)
    mesh = TetrahedronMesh.from_box([0, 1, -1, 2, 0, 0.5], nx=dims[0], ny=dims[1], nz=dims[2], device=DEV)
    node = mesh.node
    g = np.random.default_rng(3)
    if moved == "one":
        node[37, 2] += 0.01
    elif moved == "many":
        pick = torch.as_tensor(g.choice(node.shape[0], node.shape[0] // 3, replace=False), device=DEV)
        node[pick] += torch.as_tensor(0.02 * g.standard_normal((pick.numel(), 3)), device=DEV)
    space = fa.LagrangeFESpace(mesh, 1)
    assert integrators._box_grid(space, pattern_of(space)) == dims
    nc = mesh.number_of_cells()
    kc = torch.as_tensor(np.exp(g.standard_normal(nc)), device=DEV)
    for its in ([fa.ScalarDiffusionIntegrator(kc, q=2)],
                [fa.ScalarDiffusionIntegrator(kc, q=3), fa.ScalarMassIntegrator(0.5, q=3)]):
        box64 = _assemble(space, its)
        monkeypatch.setattr(integrators, "BOX_ASSEMBLY", False)
        monkeypatch.setattr(integrators, "TWO_PASS_MIN_CELLS", 0)
        two = _assemble(space, its)
        monkeypatch.setattr(integrators, "BOX_ASSEMBLY", True)
        assert float((box64 - two).abs().max()) <= 1e-15 * float(two.abs().max())
        assert torch.equal(_assemble(space, its), box64)


def test_box_axes_follow_the_nodes(monkeypatch):
    """fa_box_axes: the grid lines of a from_box mesh, None once a node
    leaves them (checked again after an in-place change of the nodes), and
    the matrix assembled through either path equals the two-pass one."""
    mesh = TetrahedronMesh.from_box([0, 2, -1, 1, 0, 3], nx=4, ny=3, nz=5, device=DEV)
    ax = mesh.box_axes()
    assert ax is not None and ax.shape[0] == 5 + 4 + 6
    ref = torch.cat([torch.linspace(0, 2, 5, dtype=torch.float64), torch.linspace(-1, 1, 4, dtype=torch.float64),
                     torch.linspace(0, 3, 6, dtype=torch.float64)])
    assert torch.allclose(ax.cpu(), ref, rtol=0, atol=1e-15)
    assert mesh.box_axes() is ax                                   # cached
    space = fa.LagrangeFESpace(mesh, 1)
    its = [fa.ScalarDiffusionIntegrator(1.5, q=2)]
    on_grid = _assemble(space, its)
    mesh.node[17, 0] += 0.05                                       # in place: the check runs again
    assert mesh.box_axes() is None
    moved = _assemble(space, its)
    monkeypatch.setattr(integrators, "BOX_ASSEMBLY", False)
    monkeypatch.setattr(integrators, "TWO_PASS_MIN_CELLS", 0)
    two = _assemble(space, its)
    assert float((moved - two).abs().max()) <= 1e-15 * float(two.abs().max())
    assert not torch.equal(moved, on_grid)


@pytest.mark.parametrize("dims", [(1, 1, 1), (4, 3, 5), (9, 7, 8)])
def test_box_load_vector_equals_two_pass(dims, monkeypatch):
    mesh = TetrahedronMesh.from_box([0, 2, -1, 1, 0, 3], nx=dims[0], ny=dims[1], nz=dims[2], device=DEV)
    space = fa.LagrangeFESpace(mesh, 1)
    nc = mesh.number_of_cells()
    g = np.random.default_rng(sum(dims) + 1)
    q = 3
    nq = space.tables(q)[1].shape[0]
    kq = torch.as_tensor(g.standard_normal((nc, nq)), device=DEV)
    kc = torch.as_tensor(np.exp(g.standard_normal(nc)), device=DEV)
    for src in (kc, kq, 2.5) + _separable_sources(kc):
        _check_load_vector(space, src, q, monkeypatch)


def _separable_sources(kc):
    from fealpy_amd.expr import cell
    return (3 * np.pi ** 2 * cell(kc) * sin(np.pi * x) * sin(np.pi * y) * sin(np.pi * z),
            sin(2 * x + 0.3) * cos(y) * exp(-0.5 * z) + 2 * x ** 2 * (y - 0.25) + 0.5 * cell(kc) * cos(3 * z) - 1.25)


def _check_load_vector(space, src, q, monkeypatch):
    form = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(src, q=q))
    b64 = form.assembly()
    b32 = form.assembly(dtype=torch.float32)
    monkeypatch.setattr(integrators, "BOX_ASSEMBLY", False)
    monkeypatch.setattr(integrators, "TWO_PASS_MIN_CELLS", 0)
    two = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(src, q=q)).assembly()
    monkeypatch.setattr(integrators, "BOX_ASSEMBLY", True)
    if isinstance(src, (torch.Tensor, float)):
        assert torch.equal(b64, two), float((b64 - two).abs().max())
    else:     # separable: the same points and factors, products possibly contracted differently
        assert float((b64 - two).abs().max()) <= 1e-15 * float(two.abs().max()), float((b64 - two).abs().max())
        assert torch.equal(form.assembly(), b64)              # run to run: bitwise
    assert torch.equal(b32, b64.to(torch.float32))


@pytest.mark.parametrize("moved", ["none", "one", "many"])
def test_box_load_vector_off_the_grid_lines(moved, monkeypatch):
    """Nodes moved off the grid lines after from_box: the cells touching
    them evaluate the separable source directly, the others from the
    tables; either way the two-pass kernels' values."""
    dims = (5, 4, 6)
    mesh = TetrahedronMesh.from_box([0, 1, 0, 2, -1, 1], nx=dims[0], ny=dims[1], nz=dims[2], device=DEV)
    node = mesh.node
    g = np.random.default_rng(7)
    if moved == "one":
        node[37, 1] += 0.01
    elif moved == "many":
        pick = torch.as_tensor(g.choice(node.shape[0], node.shape[0] // 3, replace=False), device=DEV)
        node[pick] += torch.as_tensor(0.02 * g.standard_normal((pick.numel(), 3)), device=DEV)
    space = fa.LagrangeFESpace(mesh, 1)
    assert integrators._box_grid(space, pattern_of(space)) == dims
    kc = torch.as_tensor(np.exp(g.standard_normal(mesh.number_of_cells())), device=DEV)
    for src in _separable_sources(kc):
        _check_load_vector(space, src, 3, monkeypatch)


@pytest.mark.parametrize("dims", [(1, 1, 1), (1, 6, 2), (4, 3, 5), (9, 7, 8)])
def test_box_pattern_equals_incidence_pattern(dims):
    """The grid-built CSR pattern (fa_p1_box_pattern) equals the one built
    from the incidence lists; the incidence lists, built on first use, equal
    the eager ones."""
    from fealpy_amd.fem.pattern import Pattern
    mesh = TetrahedronMesh.from_box(nx=dims[0], ny=dims[1], nz=dims[2], device=DEV)
    space = fa.LagrangeFESpace(mesh, 1)
    c2d, gdof = space.cell_to_dof(), space.number_of_global_dofs()
    box = Pattern(c2d, gdof, box=dims)
    ref = Pattern(c2d, gdof)
    assert (box.nnz, box.max_row) == (ref.nnz, ref.max_row)
    assert torch.equal(box.indptr, ref.indptr) and torch.equal(box.indices, ref.indices)
    assert box._adj is None
    assert torch.equal(box.adj_ptr, ref.adj_ptr) and torch.equal(box.adj, ref.adj)
    assert box.max_degree == ref.max_degree
    assert pattern_of(space).indices.data_ptr() != ref.indices.data_ptr()
    assert torch.equal(pattern_of(space).indices, ref.indices)


def test_box_refuses_a_foreign_pattern():
    mesh = TetrahedronMesh.from_box(nx=3, ny=3, nz=3, device=DEV)
    space = fa.LagrangeFESpace(mesh, 1)
    pat = pattern_of(space)
    _, qw, phi, _ = space.tables(2)
    L = nat.lib()
    vals = torch.zeros(pat.nnz, dtype=torch.float64, device=DEV)
    st = L.fa_p1_box_matrix(nat.FA_F64, 3, 3, 3, nat.ptr(mesh.node_rows16()), None, qw.shape[0], nat.ptr(qw),
                            nat.ptr(phi), nat.COEF_CONST, 1.0, None, nat.COEF_NONE, 0.0, None, nat.ptr(pat.indptr),
                            pat.nnz - 1, nat.ptr(vals), nat.stream())
    assert st == nat.FA_EINVAL
    assert b"not the P1 pattern" in L.fa_last_error()
    st = L.fa_p1_box_matrix(nat.FA_F64, 3, 4, 3, nat.ptr(mesh.node_rows16()), None, qw.shape[0], nat.ptr(qw),
                            nat.ptr(phi), nat.COEF_CONST, 1.0, None, nat.COEF_NONE, 0.0, None, nat.ptr(pat.indptr),
                            pat.nnz, nat.ptr(vals), nat.stream())
    assert st == nat.FA_EINVAL
    torch.cuda.synchronize()
    assert float(vals.abs().max()) == 0.0           # nothing ran


def test_threshold_mesh_keeps_the_other_kernels():
    mesh = TetrahedronMesh.from_box(nx=6, ny=6, nz=6, device=DEV, threshold=lambda p: p[..., 0] > 0.7)
    space = fa.LagrangeFESpace(mesh, 1)
    assert integrators._box_grid(space, pattern_of(space)) is None
    A = _assemble(space, [fa.ScalarDiffusionIntegrator(1.0, q=2)])
    assert torch.isfinite(A).all()
