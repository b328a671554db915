"""Host logic of the product (torch, CPU here): mesh topology, DOF numbering,
reference tables -- bit-exact against the oracle and the reference's golden
vectors; and the C-ABI library exports every symbol include/fealpy_amd.h declares."""
import os
import re

import numpy as np
import pytest
import torch

import fealpy_amd as fa
from fealpy_amd.functionspace import LagrangeFESpace
from fealpy_amd.functionspace.basis import grad_shape_function, multi_index_matrix, shape_function
from fealpy_amd.mesh import TetrahedronMesh, TriangleMesh
from fealpy_amd.quadrature import simplex_rule
from oracle import basis as ob
from oracle.mesh import tet_from_box, tri_from_box

CPU = torch.device("cpu")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("n", [(1, 1, 1), (3, 2, 1), (4, 4, 4)])
def test_tet_topology_matches_oracle(n):
    m = TetrahedronMesh.from_box(nx=n[0], ny=n[1], nz=n[2], device=CPU)
    o = tet_from_box(nx=n[0], ny=n[1], nz=n[2])
    np.testing.assert_array_equal(_np(m.cell), o.cell)
    np.testing.assert_allclose(_np(m.node), o.node, atol=0)
    for k in ("face", "edge", "face2cell", "cell2face", "cell2edge"):
        np.testing.assert_array_equal(_np(getattr(m, k)), getattr(o, k), err_msg=k)
    for p in (1, 2, 3, 4):
        np.testing.assert_array_equal(_np(m.cell_to_ipoint(p)), o.cell_to_ipoint(p), err_msg=f"c2d p={p}")
        np.testing.assert_array_equal(_np(m.face_to_ipoint(p)), o.face_to_ipoint(p), err_msg=f"f2d p={p}")
        np.testing.assert_allclose(_np(m.interpolation_points(p)), o.interpolation_points(p), atol=1e-15)
        np.testing.assert_array_equal(_np(LagrangeFESpace(m, p).is_boundary_dof()), o.is_boundary_dof(p))
        assert m.number_of_global_ipoints(p) == o.number_of_global_ipoints(p)


@pytest.mark.parametrize("n", [(1, 1), (2, 3), (5, 4)])
def test_tri_topology_matches_oracle(n):
    m = TriangleMesh.from_box(nx=n[0], ny=n[1], device=CPU)
    o = tri_from_box(nx=n[0], ny=n[1])
    np.testing.assert_array_equal(_np(m.cell), o.cell)
    for k in ("edge", "face2cell", "cell2edge"):
        np.testing.assert_array_equal(_np(getattr(m, k)), getattr(o, k), err_msg=k)
    for p in (1, 2, 3, 4):
        np.testing.assert_array_equal(_np(m.cell_to_ipoint(p)), o.cell_to_ipoint(p), err_msg=f"c2d p={p}")
        np.testing.assert_allclose(_np(m.interpolation_points(p)), o.interpolation_points(p), atol=1e-15)
        np.testing.assert_array_equal(_np(LagrangeFESpace(m, p).is_boundary_dof()), o.is_boundary_dof(p))


def test_golden_tet_dof_maps(golden):
    T = golden("ref_tetrahedron_mesh.npz")
    m = TetrahedronMesh.from_box(nx=3, ny=2, nz=1, device=CPU)
    np.testing.assert_array_equal(_np(m.face_to_ipoint(4)), T["face_to_ipoint::0::f2p"])
    np.testing.assert_array_equal(_np(m.cell_to_ipoint(4)), T["cell_to_ipoint::0::cell2ipoint"])
    np.testing.assert_allclose(_np(m.grad_lambda()), T["grad_lambda::0::glambda"], atol=1e-14)
    np.testing.assert_allclose(_np(m.entity_measure('cell')), T["entity_measure::0::cm"], atol=1e-14)
    th = lambda p: (p[..., 0] < 0.5) & (p[..., 1] < 0.5) & (p[..., 2] < 0.5)
    mt = TetrahedronMesh.from_box(nx=3, ny=2, nz=1, threshold=th, device=CPU)
    for k in ("edge", "face", "cell", "face2cell"):
        np.testing.assert_array_equal(_np(getattr(mt, k)), T["from_box::0::" + k])


def test_golden_tri_and_space(golden):
    R = golden("ref_triangle_mesh.npz")
    L = golden("ref_lagrange_fe_space.npz")
    m = TriangleMesh.from_box(nx=2, ny=2, device=CPU)
    np.testing.assert_array_equal(_np(m.cell_to_ipoint(4)), R["interpolation_point_data::0::cip"])
    np.testing.assert_allclose(_np(TriangleMesh.from_box((-1, 1, -1, 1), 2, 2, device=CPU).grad_lambda()),
                               R["grad_lambda_data::0::val"], atol=1e-14)
    s = LagrangeFESpace(TriangleMesh.from_box((0, 1, 0, 1), 1, 1, device=CPU), 2)
    np.testing.assert_array_equal(_np(s.cell_to_dof()), L["triangle_mesh_one_box::0::cell_to_dof"])
    np.testing.assert_array_equal(_np(s.is_boundary_dof()), L["triangle_mesh_one_box::0::is_boundary_dof"])


@pytest.mark.parametrize("TD,p", [(2, p) for p in (1, 2, 3, 4)] + [(3, p) for p in (1, 2, 3, 4)])
def test_reference_tables(TD, p):
    np.testing.assert_array_equal(multi_index_matrix(p, TD), ob.multi_index_matrix(p, TD))
    for q in ((1, 2, 3, 5, 9, 10, 11) if TD == 2 else (1, 2, 4, 6, 7, 8)):
        bcs, ws = simplex_rule(TD, q).get_quadrature_points_and_weights()
        ob_b, ob_w = ob.quadrature(TD, q)
        np.testing.assert_allclose(_np(bcs), ob_b, atol=1e-15)
        np.testing.assert_allclose(_np(ws), ob_w, atol=1e-15)
        np.testing.assert_allclose(shape_function(ob_b, p), ob.simplex_shape_function(ob_b, p), atol=1e-13)
        np.testing.assert_allclose(grad_shape_function(ob_b, p), ob.simplex_grad_shape_function(ob_b, p),
                                   atol=1e-12)


def test_cabi_exports_every_declared_symbol():
    """libfealpy_amd.so loads without a GPU and exports each function of include/*.h."""
    import ctypes
    from fealpy_amd import _native
    lib = _native.load()
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(fa_[a-z0-9_]+)\s*\(", src))
    assert len(names) > 10
    for n in sorted(names):
        assert hasattr(lib, n), n
        assert n in _native._SIGNATURES, f"{n} has no ctypes signature"
    assert lib.fa_version() >= 1
    assert lib.fa_pcg_work_bytes(0, 1000) > 3 * 8000


def test_product_refuses_cpu_tensors():
    """The product path has no CPU fallback: compute entry points raise."""
    from fealpy_amd import _native
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(_native.NativeError):
        _native.lib()
    mesh = fa.TetrahedronMesh.from_box(nx=1, ny=1, nz=1, device=torch.device("cpu"))
    with pytest.raises(_native.NativeError):
        mesh.bc_to_point(torch.full((1, 4), 0.25, dtype=torch.float64))


@pytest.mark.parametrize("TD,p", [(2, 2), (3, 2), (3, 3)])
def test_order_prolongation_interpolates_linears(TD, p):
    """P1 -> Pp transfer of the multigrid: exact for linear functions."""
    from fealpy_amd.solver.mg import _order_prolongation
    from fealpy_amd.sparse.tensors import CSRTensor  # noqa: F401
    mesh = (fa.TetrahedronMesh.from_box(nx=2, ny=3, nz=2, device=torch.device("cpu")) if TD == 3
            else fa.TriangleMesh.from_box(nx=3, ny=2, device=torch.device("cpu")))
    space = fa.LagrangeFESpace(mesh, p)
    rows, cols, vals = _order_prolongation(space)
    lin = lambda x: 0.3 + 1.7 * x[..., 0] - 0.4 * x[..., 1] + (2.1 * x[..., 2] if TD == 3 else 0.0)
    u1 = lin(mesh.node)
    up = torch.zeros(space.number_of_global_dofs(), dtype=torch.float64)
    up.index_add_(0, rows, vals * u1[cols])
    ref = lin(space.interpolation_points())
    assert float((up - ref).abs().max()) < 1e-14
    sums = torch.zeros_like(up).index_add_(0, rows, vals)
    assert float((sums - 1).abs().max()) < 1e-15


@pytest.mark.parametrize("dims", [(4, 6), (5, 3), (7, 7), (4, 2, 6), (3, 5, 3), (5, 5, 5), (1, 4, 3)])
def test_grid_prolongation_evaluates_the_coarse_function(dims):
    """Grid transfer of the multigrid: every fine node takes the value of the
    coarse P1 function at its position.  Checked by brute-force point location
    (barycentric coordinates of each fine node in every coarse cell) on the
    coarse from_box mesh with every count >= 2 halved, rounded up."""
    from fealpy_amd.solver.mg import _coarse_dims, _grid_prolongation
    dc = _coarse_dims(list(dims))
    M = fa.TetrahedronMesh if len(dims) == 3 else fa.TriangleMesh
    kw = dict(zip(("nx", "ny", "nz"), dims))
    kc = dict(zip(("nx", "ny", "nz"), dc))
    fine, coarse = M.from_box(**kw, device=CPU), M.from_box(**kc, device=CPU)
    v = np.random.default_rng(7).standard_normal(coarse.number_of_nodes())
    xc, cell = _np(coarse.node), _np(coarse.cell).astype(np.int64)
    xf = _np(fine.node)
    T = xc[cell[:, 1:]] - xc[cell[:, :1]]                          # (NC, d, d) edge vectors as rows
    lam = np.einsum('cij,pcj->pci', np.linalg.inv(np.transpose(T, (0, 2, 1))), xf[:, None, :] - xc[cell[:, 0]][None])
    lam = np.concatenate([1 - lam.sum(-1, keepdims=True), lam], axis=-1)   # (NP, NC, d+1)
    inside = (lam > -1e-12).all(-1)
    c = inside.argmax(1)
    assert inside.any(1).all()
    ref = (lam[np.arange(len(xf)), c] * v[cell[c]]).sum(-1)
    rows, cols, vals = _grid_prolongation(list(dims), CPU)
    got = np.zeros(len(xf))
    np.add.at(got, _np(rows), _np(vals) * v[_np(cols)])
    assert np.abs(got - ref).max() < 1e-13
    assert (_np(vals) > 0).all() and len(np.unique(_np(rows) * 10**6 + _np(cols))) == len(rows)


@pytest.mark.parametrize("prio", [False, True])
def test_block_prolongation_layout(prio):
    from fealpy_amd.solver.mg import _block_prolongation
    rows = torch.tensor([0, 1, 1, 2]); cols = torch.tensor([0, 0, 1, 1])
    vals = torch.tensor([1.0, 0.5, 0.5, 1.0], dtype=torch.float64)
    k, nf, nc = 3, 3, 2
    bd_f = torch.zeros(k * nf, dtype=torch.bool); bd_c = torch.zeros(k * nc, dtype=torch.bool)
    P = _block_prolongation(rows, cols, vals, k, prio, nf, nc, bd_f, bd_c)
    D = torch.zeros(k * nf, k * nc, dtype=torch.float64)
    r = torch.repeat_interleave(torch.arange(k * nf), (P.crow[1:] - P.crow[:-1]).long())
    D[r, P.col.long()] = P.values
    S = torch.zeros(nf, nc, dtype=torch.float64); S[rows, cols] = vals
    ref = torch.kron(torch.eye(k, dtype=torch.float64), S) if prio else torch.kron(S, torch.eye(k, dtype=torch.float64))
    assert torch.equal(D, ref)


@pytest.mark.parametrize("p", [2, 3, 4])
def test_mfma_reference_tensors_rebuild_the_gradient_product(p):
    """sum_s G_s K_s[q] == grad phi_i . grad phi_j at every point for a random
    tetrahedron (G from grad lambda_1..3), and K_6 == phi_i phi_j."""
    from fealpy_amd.fem.highorder import reference_tensors
    mesh = fa.TetrahedronMesh(torch.tensor([[0.1, 0.0, 0.2], [1.3, 0.1, -0.2], [0.2, 0.9, 0.3], [0.4, 0.3, 1.1]],
                                           dtype=torch.float64), torch.tensor([[0, 1, 2, 3]], dtype=torch.int32))
    space = fa.LagrangeFESpace(mesh, p)
    q = p + 1
    K = reference_tensors(space, q)                      # (7, NQ, ldof^2)
    gl = mesh.grad_lambda()[0]                           # (4, 3)
    pairs = ((1, 1), (2, 2), (3, 3), (1, 2), (1, 3), (2, 3))
    G = torch.tensor([float(gl[m] @ gl[n]) for m, n in pairs], dtype=torch.float64)
    got = torch.einsum('s, sqk -> qk', G, K[:6])
    bcs = space.tables(q)[0]
    gphi = mesh.grad_shape_function(bcs, p, variables='x')[0]              # (NQ, ldof, 3)
    ref = torch.einsum('qid, qjd -> qij', gphi, gphi).reshape(got.shape)
    assert float((got - ref).abs().max()) < 1e-12 * float(ref.abs().max())
    phi = space.tables(q)[2]
    assert torch.allclose(K[6], (phi[:, :, None] * phi[:, None, :]).reshape(K[6].shape), atol=0, rtol=0)


@pytest.mark.parametrize("per_lane,ks,tpg,groups", [(8, 3, 2, 7), (1, 18, 1, 13)])
def test_mfma_fragment_order(per_lane, ks, tpg, groups):
    from fealpy_amd.fem.highorder import fragment_order
    S, nq, nn = 7, ks * 2 * per_lane - 3, groups * tpg * 32 - 17
    K = torch.arange(S * nq * nn, dtype=torch.float64).reshape(S, nq, nn) + 1
    flat = fragment_order(K, ks, per_lane, tpg, groups).reshape(groups, S, tpg, ks, 64, per_lane)
    rng = np.random.default_rng(0)
    for _ in range(200):
        g, s, T, j, l, t = (int(rng.integers(0, m)) for m in (groups, S, tpg, ks, 64, per_lane))
        k = 2 * per_lane * j + per_lane * (l // 32) + t
        col = (g * tpg + T) * 32 + l % 32
        want = float(K[s, k, col]) if (k < nq and col < nn) else 0.0
        assert float(flat[g, s, T, j, l, t]) == want


def test_serving_trace_is_deterministic_and_in_range():
    from fealpy_amd.serving import make_requests, request_bytes
    a = make_requests(64, seed=7, rate=100.0)
    b = make_requests(64, seed=7, rate=100.0)
    assert [(r.nx, r.p, r.arrival) for r in a] == [(r.nx, r.p, r.arrival) for r in b]
    assert all(torch.equal(x.k, y.k) for x, y in zip(a, b))
    assert all(16 <= r.nx <= 512 and r.p in (1, 2) for r in a)
    assert all(r.k.shape[0] == 2 * r.nx * r.nx and float(r.k.min()) > 0 for r in a)
    arr = [r.arrival for r in a]
    assert arr == sorted(arr) and arr[0] > 0
    assert request_bytes(32, 2) > request_bytes(32, 1) > 0
    z = make_requests(8, seed=7)
    assert all(r.arrival == 0.0 for r in z)


def test_switched_off_packed_batches_raise():
    """While packed batches are switched off, building one raises before any
    device work (no silent fall-back); a server still takes batches (solved
    by workgroup solves)."""
    from fealpy_amd import serving
    if serving.PACKED_BATCHES_ENABLED:
        pytest.skip("packed batches are switched on")
    with pytest.raises(RuntimeError, match="switched off"):
        serving.PackedBatch([], torch.device("cpu"))
    assert serving.SolveServer(slots=1, batch=8).batch == 8


def test_sell_index_lives_with_its_pattern(monkeypatch):
    """The shared SELL index is kept on the pattern's column tensor: built
    once per pattern, reused by every matrix of it, and freed with it (no
    process-wide cache that would pin every pattern ever solved)."""
    import gc
    import weakref
    from fealpy_amd.sparse import sell
    built = []

    def fake_build(crow, col, shape):
        idx = (torch.zeros(1), torch.zeros(1), 0)
        built.append(weakref.ref(idx[0]))
        return idx
    monkeypatch.setattr(sell, "_build_index", fake_build)
    monkeypatch.setattr(torch.cuda, "current_stream", lambda *a, **k: type("S", (), {"synchronize": lambda s: None})())
    crow = torch.tensor([0, 1, 2], dtype=torch.int32)
    col = torch.tensor([0, 1], dtype=torch.int32)
    a = sell._sell_index(crow, col, (2, 2))
    assert sell._sell_index(crow, col, (2, 2)) is a and len(built) == 1
    for _ in range(20):                         # distinct patterns, dropped after use
        c2 = col.clone()
        sell._sell_index(crow, c2, (2, 2))
        del c2
    gc.collect()
    assert len(built) == 21
    assert sum(r() is not None for r in built) == 1     # only the live pattern's index remains


def test_maxit_argument_follows_the_reference():
    from fealpy_amd.solver.cg import maxit_arg
    assert maxit_arg(None) == -1 and maxit_arg(0) == 1 and maxit_arg(1) == 1 and maxit_arg(50) == 50


@pytest.mark.parametrize("dim", [2, 3])
def test_to_vtk_round_trip(tmp_path, dim):
    """to_vtk writes points, (NV, v...) connectivity, offsets, VTK types and
    the mesh's node / cell data; reading the file back gives them again."""
    from fealpy_amd.mesh.vtu import read_vtu_arrays
    cpu = torch.device("cpu")
    mesh = (fa.TriangleMesh.from_box(nx=3, ny=2, device=cpu) if dim == 2
            else fa.TetrahedronMesh.from_box(nx=2, ny=2, nz=1, device=cpu))
    nc, nn = mesh.number_of_cells(), mesh.number_of_nodes()
    mesh.nodedata['u'] = torch.arange(nn, dtype=torch.float64)
    mesh.celldata['k'] = torch.linspace(1, 2, nc, dtype=torch.float64)
    node, cl, ctype, ncl = mesh.to_vtk()
    assert ncl == nc and ctype == (5 if dim == 2 else 10)
    assert torch.equal(cl.reshape(nc, -1)[:, 1:], mesh.cell.to(cl.dtype))
    f = tmp_path / "m.vtu"
    mesh.to_vtk(fname=str(f))
    a = read_vtu_arrays(str(f))
    assert a["_counts"] == (nn, nc)
    pts = a["Points"]
    np.testing.assert_array_equal(pts[:, :dim], mesh.node.numpy())
    np.testing.assert_array_equal(a["connectivity"].reshape(nc, -1), mesh.cell.numpy())
    np.testing.assert_array_equal(a["offsets"], np.arange(1, nc + 1) * (dim + 1))
    assert np.all(a["types"] == ctype)
    np.testing.assert_array_equal(a["u"], mesh.nodedata['u'].numpy())
    np.testing.assert_array_equal(a["k"], mesh.celldata['k'].numpy())


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 1, 3), (3, 4, 2), (1, 1, 7), (5, 3, 6)])
def test_box_grid_pattern_counts_and_cell_walk(dims):
    """The structured from_box kernels' host-side facts against the oracle's
    Kuhn grid: the nonzero count and longest row of the P1 pattern
    (pattern._box_counts, which sizes the CSR arrays), and the walk the
    kernels make around a node (cubes in the order (a, b, c) = (1,1,1) ...
    (0,0,0), tetrahedra holding the node's corner in ascending t) visits
    exactly the node's cells in ascending order with the right local index,
    the vertices it derives are the cell array's, and the neighbour offsets
    in lexicographic order are the sorted CSR columns."""
    from fealpy_amd.fem.pattern import _box_counts
    nx, ny, nz = dims
    o = tet_from_box(nx=nx, ny=ny, nz=nz)
    cell = np.asarray(o.cell)
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    cols = [set() for _ in range(nn)]
    inc = [[] for _ in range(nn)]
    for c in range(cell.shape[0]):
        for i in range(4):
            inc[cell[c, i]].append((c, i))
            cols[cell[c, i]].update(int(v) for v in cell[c])
    assert _box_counts(nx, ny, nz) == (sum(map(len, cols)), max(map(len, cols)))
    corner = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    kuhn = [[0, 1, 2, 6], [0, 5, 1, 6], [0, 4, 5, 6], [0, 7, 4, 6], [0, 3, 7, 6], [0, 2, 3, 6]]
    nz1, nyz = nz + 1, (ny + 1) * (nz + 1)
    for r in range(nn):
        ix, iy, iz = r // nyz, (r % nyz) // nz1, r % nz1
        walk = []
        for a, b, c in [(1, 1, 1), (1, 1, 0), (1, 0, 1), (1, 0, 0), (0, 1, 1), (0, 1, 0), (0, 0, 1), (0, 0, 0)]:
            cx, cy, cz = ix - a, iy - b, iz - c
            if not (0 <= cx < nx and 0 <= cy < ny and 0 <= cz < nz):
                continue
            kap = corner.index((a, b, c))
            for t in range(6):
                if kap in kuhn[t]:
                    cid = 6 * ((cx * ny + cy) * nz + cz) + t
                    verts = [(cx + corner[k][0]) * nyz + (cy + corner[k][1]) * nz1 + cz + corner[k][2]
                             for k in kuhn[t]]
                    assert list(cell[cid]) == verts
                    walk.append((cid, kuhn[t].index(kap)))
        assert walk == inc[r]
        offs = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)
                if (min(dx, dy, dz) >= 0 or max(dx, dy, dz) <= 0)
                and 0 <= ix + dx <= nx and 0 <= iy + dy <= ny and 0 <= iz + dz <= nz]
        assert [r + dx * nyz + dy * nz1 + dz for dx, dy, dz in offs] == sorted(cols[r])


@pytest.mark.parametrize("cls,dims", [("tet", (3, 2, 4)), ("tri", (5, 3))])
def test_topology_on_first_use(cls, dims):
    """Faces, edges and the from_box side tags are built on first use: a P1
    count and the P1 boundary of a full box grid (from grid positions) need
    none of them, and equal what the face numbering gives; the tags appear
    on any dict access."""
    if cls == "tet":
        mesh = TetrahedronMesh.from_box(nx=dims[0], ny=dims[1], nz=dims[2], device="cpu")
    else:
        mesh = TriangleMesh.from_box([0, 1, 0, 1], nx=dims[0], ny=dims[1], device="cpu")
    assert "_constructed" not in mesh.__dict__
    assert mesh.number_of_global_ipoints(1) == mesh.number_of_nodes()
    flag = mesh.is_boundary_dof(1)
    assert "_constructed" not in mesh.__dict__
    ref = torch.zeros_like(flag)
    ref[mesh.face[mesh.boundary_face_flag()].reshape(-1).long()] = True      # builds the topology
    assert mesh.__dict__["_constructed"] and torch.equal(flag, ref)
    if cls == "tet":
        assert "upface" in mesh.meshdata and set(mesh.meshdata) >= {"from_box", "leftface", "bottomface"}
        tags = torch.cat([mesh.meshdata[k] for k in ("leftface", "rightface", "frontface", "backface",
                                                     "bottomface", "upface")])
        assert torch.equal(torch.sort(tags).values, mesh.boundary_face_index())
    with pytest.raises(AttributeError):
        mesh.no_such_attribute
