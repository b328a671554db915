"""Simplex meshes, topology and Lagrange DOF numbering (NumPy, FP64/int32).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""
import numpy as np

from .basis import multi_index_matrix, quadrature, simplex_grad_shape_function, simplex_shape_function


def flocc(array):
    """fealpy/mesh/utils.py:81-110 -- first/last occurrence of each unique row
    (lexicographic order, column 0 primary) and the original->unique map."""
    indices = np.lexsort(tuple(reversed(array.T)), axis=0)
    sorted_array = array[indices]
    diff = np.any(sorted_array[1:] != sorted_array[:-1], axis=1)
    diff = np.concatenate([[True], diff, [True]])
    group_index = np.cumsum(diff[:-1]) - 1
    i0 = indices[diff[:-1]]
    i1 = indices[diff[1:]]
    j = np.empty_like(indices)
    j[indices] = np.arange(len(indices))
    return i0, i1, group_index[j]


def simplex_gdof(p, nums):
    """fealpy/mesh/utils.py:165-173."""
    coef, count = 1, nums[0]
    for i in range(1, len(nums)):
        coef = (coef * (p - i)) // i
        count += coef * nums[i]
    return count


class SimplexMesh:
    """Triangle (TD=2) or tetrahedron (TD=3) mesh with the reference topology."""

    TRI_LOCAL_EDGE = np.array([(1, 2), (2, 0), (0, 1)])                      # triangle_mesh.py:24-25
    TET_LOCAL_EDGE = np.array([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)])  # tetrahedron_mesh.py:21-22
    TET_LOCAL_FACE = np.array([(1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1)])      # tetrahedron_mesh.py:23-24
    TET_LOCAL_FACE2EDGE = np.array([(5, 4, 3), (5, 1, 2), (4, 2, 0), (3, 0, 1)])  # tetrahedron_mesh.py:37-38

    def __init__(self, node, cell):
        self.node = np.asarray(node, dtype=np.float64)
        self.cell = np.asarray(cell)
        self.TD = self.cell.shape[1] - 1
        self.GD = self.node.shape[1]
        self.itype = self.cell.dtype
        self.construct()

    # -- topology: fealpy/mesh/mesh_data_structure.py:428-464 ---------------
    def construct(self):
        cell, TD = self.cell, self.TD
        local_face = self.TRI_LOCAL_EDGE if TD == 2 else self.TET_LOCAL_FACE
        NC = cell.shape[0]
        NFC = local_face.shape[0]
        total_face = cell[:, local_face].reshape(-1, local_face.shape[1])
        i0, i1, j = flocc(np.sort(total_face, axis=1))
        self.face = total_face[i0, :]
        self.cell2face = j.reshape(NC, NFC).astype(self.itype)
        self.face2cell = np.stack([i0 // NFC, i1 // NFC, i0 % NFC, i1 % NFC], axis=-1).astype(self.itype)
        if TD == 3:
            total_edge = cell[:, self.TET_LOCAL_EDGE].reshape(-1, 2)
            i2, _, j = flocc(np.sort(total_edge, axis=1))
            self.edge = total_edge[i2, :]
            self.cell2edge = j.reshape(NC, 6).astype(self.itype)
        else:
            self.edge = self.face
            self.cell2edge = self.cell2face

    def number_of_nodes(self): return self.node.shape[0]
    def number_of_edges(self): return self.edge.shape[0]
    def number_of_faces(self): return self.face.shape[0]
    def number_of_cells(self): return self.cell.shape[0]

    def boundary_face_flag(self):
        """mesh_data_structure.py:368-374."""
        return self.face2cell[:, 0] == self.face2cell[:, 1]

    def boundary_face_index(self):
        return np.nonzero(self.boundary_face_flag())[0]

    def face_to_edge(self):
        """mesh_data_structure.py:219-226 (tet)."""
        return self.cell2edge[self.face2cell[:, [0]], self.TET_LOCAL_FACE2EDGE[self.face2cell[:, 2]]]

    # -- geometry -----------------------------------------------------------
    def simplex_measure(self):
        """numpy_backend.py:413-421 (det of consecutive edge vectors / TD!)."""
        points = self.node[self.cell]
        edges = points[..., 1:, :] - points[..., :-1, :]
        return np.linalg.det(edges) / (2.0 if self.TD == 2 else 6.0)

    def entity_measure(self, etype='cell'):
        """triangle_mesh.py:45-70 / tetrahedron_mesh.py:165-175 (cell), :186-195 (face)."""
        node = self.node
        if etype in ('cell', self.TD):
            if self.TD == 2:
                return self.simplex_measure()
            c = self.cell
            v01 = node[c[:, 1]] - node[c[:, 0]]
            v02 = node[c[:, 2]] - node[c[:, 0]]
            v03 = node[c[:, 3]] - node[c[:, 0]]
            return np.sum(v03 * np.cross(v01, v02), axis=1) / 6.0
        if etype in ('edge', 1):
            e = self.edge
            return np.linalg.norm(node[e[:, 0]] - node[e[:, 1]], axis=-1)
        if etype in ('face', 2) and self.TD == 3:
            f = self.face
            nv = np.cross(node[f[:, 1]] - node[f[:, 0]], node[f[:, 2]] - node[f[:, 0]])
            return np.sqrt(np.square(nv).sum(axis=1)) / 2.0
        if etype in ('node', 0):
            return np.zeros(1)
        raise ValueError(etype)

    def grad_lambda(self):
        """Triangle: numpy_backend.py:586-598; tetrahedron: numpy_backend.py:619-629."""
        node, cell = self.node, self.cell
        if self.TD == 2:
            pts = node[cell]
            e0 = pts[..., 2, :] - pts[..., 1, :]
            e1 = pts[..., 0, :] - pts[..., 2, :]
            e2 = pts[..., 1, :] - pts[..., 0, :]
            nv = np.linalg.det(np.stack([e0, e1], axis=-2))
            res = np.stack([np.flip(e0, -1), np.flip(e1, -1), np.flip(e2, -1)], axis=-2)
            res[..., 0] *= -1
            return res / nv[:, None, None]
        NC = cell.shape[0]
        out = np.zeros((NC, 4, 3))
        vol = self.simplex_measure()
        for i in range(4):
            j, k, m = self.TET_LOCAL_FACE[i]
            vjk = node[cell[:, k]] - node[cell[:, j]]
            vjm = node[cell[:, m]] - node[cell[:, j]]
            out[:, i, :] = np.cross(vjm, vjk) / (6 * vol.reshape(-1, 1))
        return out

    def bc_to_point(self, bcs):
        """numpy_backend.py:401-407 -> (NC, NQ, GD)."""
        return np.einsum('ijk, ...j -> i...k', self.node[self.cell], bcs)

    def quadrature_formula(self, q):
        return quadrature(self.TD, q)

    def shape_function(self, bcs, p):
        return simplex_shape_function(bcs, p, multi_index_matrix(p, self.TD))

    def grad_shape_function(self, bcs, p, variables='x'):
        """mesh_base.py:713-749 (triangle override triangle_mesh.py:142-155)."""
        R = simplex_grad_shape_function(bcs, p, multi_index_matrix(p, self.TD))
        if variables == 'u':
            return R
        return np.einsum('...bm, qjb -> ...qjm', self.grad_lambda(), R)

    # -- interpolation points / DOF numbering --------------------------------
    def number_of_global_ipoints(self, p):
        if self.TD == 2:
            return simplex_gdof(p, (self.number_of_nodes(), self.number_of_edges(), self.number_of_cells()))
        return simplex_gdof(p, (self.number_of_nodes(), self.number_of_edges(),
                                self.number_of_faces(), self.number_of_cells()))

    def number_of_local_ipoints(self, p, iptype='cell'):
        d = {'cell': self.TD, 'face': self.TD - 1, 'edge': 1, 'node': 0}.get(iptype, iptype)
        from math import comb
        return 1 if d == 0 else comb(p + d, d)

    def edge_to_ipoint(self, p):
        """mesh_base.py:188-211."""
        NN, NE = self.number_of_nodes(), self.number_of_edges()
        e = self.edge
        idx = np.arange(NE, dtype=e.dtype)
        return np.concatenate([e[:, 0:1], (p - 1) * idx.reshape(-1, 1) + np.arange(0, p - 1, dtype=e.dtype) + NN,
                               e[:, 1:2]], axis=-1)

    def face_to_ipoint(self, p):
        """Triangle: = edge_to_ipoint (triangle_mesh.py:272-273); tet: tetrahedron_mesh.py:348-386."""
        if self.TD == 2:
            return self.edge_to_ipoint(p)
        fdof = (p + 1) * (p + 2) // 2
        edgeIdx = np.zeros((2, p + 1), dtype=np.int64)
        edgeIdx[0] = np.arange(p + 1)
        edgeIdx[1] = edgeIdx[0][::-1]
        NN, NE, NF = self.number_of_nodes(), self.number_of_edges(), self.number_of_faces()
        face, edge = self.face, self.edge
        face2edge = self.face_to_edge()
        e2p = self.edge_to_ipoint(p)
        f2p = np.zeros((NF, fdof), dtype=self.itype)
        isEdgeIPoint = multi_index_matrix(p, 2) == 0
        fe = np.array([1, 0, 0])
        for i in range(3):
            I = np.ones(NF, dtype=np.int64)
            sign = face[:, fe[i]] == edge[face2edge[:, i], 0]
            I[sign] = 0
            f2p[:, isEdgeIPoint[:, i]] = e2p[face2edge[:, [i]], edgeIdx[I]]
        if p > 2:
            base = NN + (p - 1) * NE
            inner = ~(isEdgeIPoint[:, 0] | isEdgeIPoint[:, 1] | isEdgeIPoint[:, 2])
            fidof = fdof - 3 * p
            f2p[:, inner] = base + np.arange(NF * fidof, dtype=self.itype).reshape(NF, fidof)
        return f2p

    def cell_to_ipoint(self, p):
        """Triangle: triangle_mesh.py:218-270; tetrahedron: tetrahedron_mesh.py:388-441."""
        cell = self.cell
        if p == 1:
            return cell.copy()
        NN, NE, NC = self.number_of_nodes(), self.number_of_edges(), self.number_of_cells()
        if self.TD == 2:
            mi = multi_index_matrix(p, 2)
            idx0 = np.nonzero(mi[:, 0] == 0)[0]
            idx1 = np.nonzero(mi[:, 1] == 0)[0]
            idx2 = np.nonzero(mi[:, 2] == 0)[0]
            f2c = self.face2cell
            e2p = self.edge_to_ipoint(p)
            ldof = (p + 1) * (p + 2) // 2
            c2p = np.zeros((NC, ldof), dtype=cell.dtype)
            flag = f2c[:, 2] == 0
            c2p[f2c[flag, 0][:, None], idx0] = e2p[flag]
            flag = f2c[:, 2] == 1
            c2p[f2c[flag, 0][:, None], idx1[::-1]] = e2p[flag]
            flag = f2c[:, 2] == 2
            c2p[f2c[flag, 0][:, None], idx2] = e2p[flag]
            iflag = f2c[:, 0] != f2c[:, 1]
            flag = iflag & (f2c[:, 3] == 0)
            c2p[f2c[flag, 1][:, None], idx0[::-1]] = e2p[flag]
            flag = iflag & (f2c[:, 3] == 1)
            c2p[f2c[flag, 1][:, None], idx1] = e2p[flag]
            flag = iflag & (f2c[:, 3] == 2)
            c2p[f2c[flag, 1][:, None], idx2[::-1]] = e2p[flag]
            cdof = (p - 1) * (p - 2) // 2
            flag = np.sum(mi > 0, axis=1) == 3
            c2p[:, flag] = NN + NE * (p - 1) + np.arange(NC * cdof, dtype=cell.dtype).reshape(NC, cdof)
            return c2p
        # tetrahedron
        NF = self.number_of_faces()
        fdof = (p + 1) * (p + 2) // 2
        ldof = (p + 1) * (p + 2) * (p + 3) // 6
        face, c2f = self.face, self.cell2face
        c2p = np.zeros((NC, ldof), dtype=self.itype)
        f2p = self.face_to_ipoint(p)
        m2 = multi_index_matrix(p, 2).T
        m3 = multi_index_matrix(p, 3).T
        isFaceIPoint = m3 == 0
        fidx = np.argsort(np.argsort(face, axis=1), axis=1)
        for i in range(4):
            idx = [k for k in range(4) if k != i]
            idxj = np.argsort(cell[:, idx], axis=1)
            idxi = fidx[c2f[:, i]]
            order = idxj[np.arange(NC).reshape(-1, 1), idxi]
            mi = m2[order]
            k = mi[:, 1] + mi[:, 2]
            a = k * (k + 1) // 2 + mi[:, 2]
            c2p[:, isFaceIPoint[i]] = f2p[c2f[:, [i]], a]
        if p > 3:
            base = NN + (p - 1) * NE + (fdof - 3 * p) * NF
            idof = ldof - 4 - 6 * (p - 1) - 4 * (fdof - 3 * p)
            inner = ~(isFaceIPoint[0] | isFaceIPoint[1] | isFaceIPoint[2] | isFaceIPoint[3])
            c2p[:, inner] = base + np.arange(NC * idof, dtype=self.itype).reshape(NC, idof)
        return c2p

    def interpolation_points(self, p):
        """triangle_mesh.py:180-216 / tetrahedron_mesh.py:302-346."""
        node = self.node
        if p == 1:
            return node.copy()
        GD = self.GD
        parts = [node]
        w = multi_index_matrix(p, 1).astype(np.float64)[1:-1] / p
        parts.append(np.einsum('ij, ...jm->...im', w, node[self.edge]).reshape(-1, GD))
        if self.TD == 2:
            if p >= 3:
                mi = multi_index_matrix(p, 2).astype(np.float64)
                inner = np.all(mi > 0, axis=1)
                parts.append(np.einsum('ij, kj...->ki...', mi[inner] / p, node[self.cell]).reshape(-1, GD))
            return np.concatenate(parts, axis=0)
        if p > 2:
            mi = multi_index_matrix(p, 2).astype(np.float64)
            inner = np.sum(mi > 0, axis=-1) == 3
            parts.append(np.einsum('ij, kj...->ki...', mi[inner] / p, node[self.face]).reshape(-1, GD))
        if p > 3:
            mi = multi_index_matrix(p, 3).astype(np.float64)
            inner = np.sum(mi > 0, axis=-1) == 4
            parts.append(np.einsum('ij, kj...->ki...', mi[inner] / p, node[self.cell]).reshape(-1, GD))
        return np.concatenate(parts, axis=0)

    def is_boundary_dof(self, p):
        """functionspace/dofs.py:23-55 (threshold None, method None/'interp')."""
        gdof = self.number_of_global_ipoints(p)
        f2d = self.face_to_ipoint(p)[self.boundary_face_index()]
        flag = np.zeros(gdof, dtype=bool)
        flag[f2d] = True
        return flag


def tri_from_box(box=(0, 1, 0, 1), nx=10, ny=10):
    """fealpy/mesh/triangle_mesh.py:1387-1436 (no threshold)."""
    NN = (nx + 1) * (ny + 1)
    x = np.linspace(box[0], box[1], nx + 1)
    y = np.linspace(box[2], box[3], ny + 1)
    X, Y = np.meshgrid(x, y, indexing='ij')
    node = np.concatenate((X.reshape(-1, 1), Y.reshape(-1, 1)), axis=1)
    idx = np.arange(NN, dtype=np.int32).reshape(nx + 1, ny + 1)
    cell0 = np.concatenate((idx[1:, 0:-1].T.reshape(-1, 1), idx[1:, 1:].T.reshape(-1, 1),
                            idx[0:-1, 0:-1].T.reshape(-1, 1)), axis=1)
    cell1 = np.concatenate((idx[0:-1, 1:].T.reshape(-1, 1), idx[0:-1, 0:-1].T.reshape(-1, 1),
                            idx[1:, 1:].T.reshape(-1, 1)), axis=1)
    return SimplexMesh(node, np.concatenate((cell0, cell1), axis=0))


def tet_from_box(box=(0, 1, 0, 1, 0, 1), nx=10, ny=10, nz=10, threshold=None):
    """fealpy/mesh/tetrahedron_mesh.py:1017-1086 (Kuhn: 6 tets per cube around the main diagonal)."""
    NN = (nx + 1) * (ny + 1) * (nz + 1)
    x = np.linspace(box[0], box[1], nx + 1)
    y = np.linspace(box[2], box[3], ny + 1)
    z = np.linspace(box[4], box[5], nz + 1)
    X, Y, Z = np.meshgrid(x, y, z, indexing='ij')
    node = np.concatenate((X.reshape(-1, 1), Y.reshape(-1, 1), Z.reshape(-1, 1)), axis=1)
    idx = np.arange(NN, dtype=np.int32).reshape(nx + 1, ny + 1, nz + 1)
    nyz = (ny + 1) * (nz + 1)
    c0 = idx[:-1, :-1, :-1]
    c1 = c0 + nyz
    c2 = c1 + nz + 1
    c3 = c0 + nz + 1
    c4 = c0 + 1
    c5 = c4 + nyz
    c6 = c5 + nz + 1
    c7 = c4 + nz + 1
    cell = np.concatenate([c.reshape(-1, 1) for c in (c0, c1, c2, c3, c4, c5, c6, c7)], axis=1)
    local = np.array([[0, 1, 2, 6], [0, 5, 1, 6], [0, 4, 5, 6],
                      [0, 7, 4, 6], [0, 3, 7, 6], [0, 2, 3, 6]], dtype=np.int32)
    cell = cell[:, local].reshape(-1, 4)
    if threshold is not None:
        bc = np.sum(node[cell, :], axis=1) / cell.shape[1]
        cell = cell[~threshold(bc)]
        valid = np.zeros(NN, dtype=bool)
        valid[cell] = True
        node = node[valid]
        idxmap = np.zeros(NN, dtype=cell.dtype)
        idxmap[valid] = np.arange(valid.sum(), dtype=cell.dtype)
        cell = idxmap[cell]
    return SimplexMesh(node, cell)
