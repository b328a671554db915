"""Device mesh topology (fa_topology_sort / fa_topology_fill) against the
oracle's construct + flocc (pinned to the reference's mesh data), bit-exact:
faces, edges, face -> cell, cell -> face, cell -> edge and the Lagrange DOF
maps built on them.  Also a mesh with its cells shuffled and reoriented,
where no from_box structure is left, against the oracle on the same cells."""
import numpy as np
import pytest
import torch

from fealpy_amd.functionspace import LagrangeFESpace
from fealpy_amd.mesh import TetrahedronMesh, TriangleMesh
from oracle.mesh import SimplexMesh as OSimplex
from oracle.mesh import tet_from_box, tri_from_box

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("n", [(1, 1, 1), (3, 2, 1), (4, 4, 4), (9, 7, 5)])
def test_tet_topology_device(n):
    m = TetrahedronMesh.from_box(nx=n[0], ny=n[1], nz=n[2], device=DEV)
    o = tet_from_box(nx=n[0], ny=n[1], nz=n[2])
    for k in ("face", "edge", "face2cell", "cell2face", "cell2edge"):
        np.testing.assert_array_equal(_np(getattr(m, k)), getattr(o, k), err_msg=k)
    for p in (2, 3, 4):
        np.testing.assert_array_equal(_np(m.cell_to_ipoint(p)), o.cell_to_ipoint(p), err_msg=f"c2d p={p}")
        np.testing.assert_array_equal(_np(LagrangeFESpace(m, p).is_boundary_dof()), o.is_boundary_dof(p))


@pytest.mark.parametrize("n", [(1, 1), (2, 3), (17, 11)])
def test_tri_topology_device(n):
    m = TriangleMesh.from_box(nx=n[0], ny=n[1], device=DEV)
    o = tri_from_box(nx=n[0], ny=n[1])
    for k in ("edge", "face2cell", "cell2edge"):
        np.testing.assert_array_equal(_np(getattr(m, k)), getattr(o, k), err_msg=k)
    for p in (2, 3):
        np.testing.assert_array_equal(_np(m.cell_to_ipoint(p)), o.cell_to_ipoint(p), err_msg=f"c2d p={p}")


def test_shuffled_cells_device():
    """Cells permuted and their vertices rotated: the first occurrence of a
    face is then often not in the lowest-numbered cell of the box order."""
    o0 = tet_from_box(nx=6, ny=5, nz=4)
    rng = np.random.default_rng(2)
    cell = o0.cell[rng.permutation(o0.cell.shape[0])]
    # an even permutation of each cell's vertices keeps the orientation
    rot = np.array([[0, 1, 2, 3], [1, 2, 0, 3], [2, 0, 1, 3]])[rng.integers(0, 3, cell.shape[0])]
    cell = np.take_along_axis(cell, rot, axis=1).astype(np.int32)
    o = OSimplex(o0.node, cell)
    m = TetrahedronMesh(torch.as_tensor(o0.node, device=DEV), torch.as_tensor(cell, device=DEV))
    for k in ("face", "edge", "face2cell", "cell2face", "cell2edge"):
        np.testing.assert_array_equal(_np(getattr(m, k)), getattr(o, k), err_msg=k)
