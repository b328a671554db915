"""Triangle and tetrahedron meshes on torch tensors.

Mesh generation and topology are set-up work outside the assemble-and-solve
hot path; they run as torch tensor programs on the mesh's device (the GPU by
default).  The numbering they produce is the reference's, because every
global DOF index downstream depends on it:

* faces/edges are the unique sorted vertex tuples in lexicographic order,
  stored with the orientation of their first occurrence, face2cell =
  [c0, c1, l0, l1] from the first/last occurrence
  (fealpy/mesh/mesh_data_structure.py:428-464, fealpy/mesh/utils.py:81-110);
* Lagrange DOFs: nodes, then p-1 per edge, then face-interior, then
  cell-interior points; local order = multi-index order
  (triangle_mesh.py:218-270, tetrahedron_mesh.py:348-441, mesh_base.py:188-211).
"""
import ctypes
import threading
from math import comb

import numpy as np
import torch

from ..functionspace.basis import grad_shape_function as _grad_shape, multi_index_matrix, \
    shape_function as _shape
from ..quadrature import simplex_rule

_S = slice(None)


def default_device():
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


def _lexsort_rows(a):
    """Stable permutation sorting the rows of an int tensor lexicographically
    (column 0 most significant).  Columns are packed into int64 keys and sorted
    least-significant pass first with stable sorts."""
    n, m = a.shape
    a = a.to(torch.int64)
    span = int(a.max().item()) + 1 if n else 1
    per = 1
    while per < m and span ** (per + 1) < (1 << 62):
        per += 1
    idx = torch.arange(n, device=a.device)
    hi = m
    while hi > 0:
        lo = max(0, hi - per)
        key = torch.zeros(n, dtype=torch.int64, device=a.device)
        for k in range(lo, hi):
            key = key * span + a[idx, k]
        _, o = torch.sort(key, stable=True)
        idx = idx[o]
        hi = lo
    return idx


def flocc(array):
    """First occurrence, last occurrence and original->unique map of the rows."""
    idx = _lexsort_rows(array)
    s = array[idx]
    diff = torch.any(s[1:] != s[:-1], dim=1)
    one = torch.ones(1, dtype=torch.bool, device=array.device)
    flag = torch.cat([one, diff, one])
    group = torch.cumsum(flag[:-1].to(torch.int64), 0) - 1
    i0 = idx[flag[:-1]]
    i1 = idx[flag[1:]]
    j = torch.empty_like(idx)
    j[idx] = group
    return i0, i1, j


def _device_entities(cell, local, nn):
    """(entity, cell -> entity, entity -> cell) of the local entities `local`
    (nloc x m vertex positions) of every cell, by the device topology kernels
    (fa_topology_sort / fa_topology_fill): the numbering of flocc above."""
    from .. import _native as nat
    L = nat.lib()
    dev = cell.device
    cell = cell.to(torch.int32).contiguous()
    loc = local.to(device=dev, dtype=torch.int32).contiguous()
    nc, nv = cell.shape
    nloc, m = loc.shape
    ws = nat.workspace(L.fa_topology_work_bytes(nc * nloc, nn), dev)
    nu = nat.ctypes.c_int64(0)
    nat.check(L.fa_topology_sort(m, nloc, nat.ptr(loc), nv, nat.ptr(cell), nc, nn, nat.ctypes.byref(nu),
                                 nat.ptr(ws), ws.numel(), nat.stream()), "fa_topology_sort")
    n = int(nu.value)
    entity = torch.empty((n, m), dtype=torch.int32, device=dev)
    c2e = torch.empty((nc, nloc), dtype=torch.int32, device=dev)
    e2c = torch.empty((n, 4), dtype=torch.int32, device=dev)
    nat.check(L.fa_topology_fill(m, nloc, nat.ptr(loc), nv, nat.ptr(cell), nc, nn, nat.ptr(entity), nat.ptr(c2e),
                                 nat.ptr(e2c), nat.ptr(ws), ws.numel(), nat.stream()), "fa_topology_fill")
    return entity, c2e, e2c


def _gdof(p, nums):
    """Number of Lagrange DOFs of degree p from entity counts (nodes, edges[, faces], cells)."""
    total = nums[0]
    for d in range(1, len(nums)):
        total += comb(p - 1, d) * nums[d]
    return total


_TOPOLOGY = frozenset(('face', 'edge', 'cell2face', 'face2cell', 'cell2edge', 'edge2cell'))
_BOX_FACE_TAGS = ('leftface', 'rightface', 'frontface', 'backface', 'bottomface', 'upface')


class _MeshData(dict):
    """mesh.meshdata: a dict whose from_box face tags (the boundary faces of
    each side of the box, which need the face numbering) are computed on
    first access by any dict operation, so a mesh that never asks for them
    never builds its faces."""

    def __init__(self):
        super().__init__()
        self._tagger = None

    def _fill(self):
        tag, self._tagger = self._tagger, None
        if tag is not None:
            tag()

    def __missing__(self, key):
        if key in _BOX_FACE_TAGS and self._tagger is not None:
            self._fill()
            return dict.__getitem__(self, key)
        raise KeyError(key)

    def get(self, key, default=None):
        if key in _BOX_FACE_TAGS:
            self._fill()
        return dict.get(self, key, default)

    def __contains__(self, key):
        if key in _BOX_FACE_TAGS:
            self._fill()
        return dict.__contains__(self, key)

    def keys(self):
        self._fill()
        return dict.keys(self)

    def items(self):
        self._fill()
        return dict.items(self)

    def values(self):
        self._fill()
        return dict.values(self)

    def __iter__(self):
        self._fill()
        return dict.__iter__(self)

    def __len__(self):
        self._fill()
        return dict.__len__(self)

    def __repr__(self):
        self._fill()
        return dict.__repr__(self)


class SimplexMesh:
    TD = None
    localEdge = None
    localFace = None

    def __init__(self, node, cell):
        if not torch.is_tensor(node):
            node = torch.as_tensor(np.asarray(node), dtype=torch.float64)
        if not torch.is_tensor(cell):
            cell = torch.as_tensor(np.asarray(cell), dtype=torch.int32)
        self.node = node
        self.cell = cell
        self.device = node.device
        self.itype = cell.dtype
        self.ftype = node.dtype
        self.GD = node.shape[1]
        self.nodedata, self.edgedata, self.facedata, self.celldata = {}, {}, {}, {}
        self.meshdata = _MeshData()
        self._cache = {}
        self._topo_lock = threading.Lock()

    def __getattr__(self, name):
        # faces, edges and their cell maps are built on first use (construct);
        # a P1 solve on a from_box grid needs none of them
        d = self.__dict__
        if name in _TOPOLOGY and '_topo_lock' in d:
            with d['_topo_lock']:
                if not d.get('_constructed', False):
                    self.construct()
                    d['_constructed'] = True
                    if self.cell.is_cuda:   # complete before another stream or thread reads it
                        torch.cuda.current_stream(self.cell.device).synchronize()
            if name in d:
                return d[name]
        raise AttributeError(f"{type(self).__name__!r} object has no attribute {name!r}")

    def node_rows16(self):
        """Vertex coordinates as 16-byte-aligned rows for wide loads in the
        assembly kernels: the node array itself in 2D ((NN, 2) FP64), a copy
        padded to (NN, 4) in 3D, refreshed when `node` is replaced or changed
        in place."""
        node = self.node
        if self.GD == 2 and node.dtype == torch.float64 and node.is_contiguous():
            return node
        key = (id(node), node._version)
        hit = self._cache.get('_node_rows16')
        if hit is not None and hit[0] == key and hit[1] is node:
            return hit[2]
        wide = torch.zeros((node.shape[0], 4), dtype=torch.float64, device=node.device)
        wide[:, :node.shape[1]] = node
        self._cache['_node_rows16'] = (key, node, wide)
        return wide

    def box_axes(self):
        """The grid lines (X, Y, Z concatenated, FP64 on the device) of a 3D
        from_box mesh whose every node sits on them (fa_box_axes), else None;
        checked once per node array and its in-place changes."""
        grid = self._box_grid()
        if grid is None or self.GD != 3:
            return None
        node = self.node
        key = (id(node), node._version)
        hit = self._cache.get('_box_axes')
        if hit is not None and hit[0] == key and hit[1] is node:
            return hit[2]
        from .. import _native as nat
        L = nat.lib()
        nx, ny, nz = grid
        axes = torch.empty(nx + ny + nz + 3, dtype=torch.float64, device=node.device)
        ws = nat.workspace(L.fa_box_axes_work_bytes(), node.device)
        on = ctypes.c_int32(0)
        nat.check(L.fa_box_axes(nx, ny, nz, nat.ptr(self.node_rows16()), nat.ptr(axes), ctypes.byref(on),
                                nat.ptr(ws), ws.numel(), nat.stream()), "fa_box_axes")
        out = axes if on.value else None
        self._cache['_box_axes'] = (key, node, out)
        return out

    # -- topology ---------------------------------------------------------------
    def construct(self):
        self.__dict__['_constructed'] = True
        cell = self.cell
        NC = cell.shape[0]
        if cell.is_cuda and NC > 0:
            it = self.itype
            self.face, self.cell2face, self.face2cell = (t.to(it) for t in
                                                         _device_entities(cell, self.localFace, self.node.shape[0]))
            if self.TD == 3:
                self.edge, self.cell2edge = (t.to(it) for t in
                                             _device_entities(cell, self.localEdge, self.node.shape[0])[:2])
            else:
                self.edge = self.face
                self.cell2edge = self.cell2face
                self.edge2cell = self.face2cell
            return
        lf = self.localFace.to(self.device)
        NFC = lf.shape[0]
        total_face = cell[:, lf].reshape(-1, lf.shape[1])
        i0, i1, j = flocc(torch.sort(total_face, dim=1).values)
        self.face = total_face[i0]
        self.cell2face = j.reshape(NC, NFC).to(self.itype)
        self.face2cell = torch.stack([i0 // NFC, i1 // NFC, i0 % NFC, i1 % NFC], dim=-1).to(self.itype)
        if self.TD == 3:
            le = self.localEdge.to(self.device)
            total_edge = cell[:, le].reshape(-1, 2)
            i2, _, j2 = flocc(torch.sort(total_edge, dim=1).values)
            self.edge = total_edge[i2]
            self.cell2edge = j2.reshape(NC, le.shape[0]).to(self.itype)
        else:
            self.edge = self.face
            self.cell2edge = self.cell2face
            self.edge2cell = self.face2cell

    def top_dimension(self):
        return self.TD

    def geo_dimension(self):
        return self.GD

    def number_of_nodes(self):
        return self.node.shape[0]

    def number_of_edges(self):
        return self.edge.shape[0]

    def number_of_faces(self):
        return self.face.shape[0]

    def number_of_cells(self):
        return self.cell.shape[0]

    def count(self, etype):
        return {'cell': self.number_of_cells(), 'face': self.number_of_faces(), 'edge': self.number_of_edges(),
                'node': self.number_of_nodes(), self.TD: self.number_of_cells(), 0: self.number_of_nodes(),
                1: self.number_of_edges(), self.TD - 1: self.number_of_faces()}[etype]

    def entity(self, etype, index=_S):
        name = {0: 'node', 1: 'edge', self.TD - 1: 'face', self.TD: 'cell'}.get(etype, etype)
        return getattr(self, name)[index]

    def face_to_cell(self, index=_S):
        return self.face2cell[index]

    def cell_to_face(self, index=_S):
        return self.cell2face[index]

    def cell_to_edge(self, index=_S):
        return self.cell2edge[index]

    def boundary_face_flag(self):
        return self.face2cell[:, 0] == self.face2cell[:, 1]

    def boundary_face_index(self):
        return torch.nonzero(self.boundary_face_flag())[:, 0]

    def boundary_node_flag(self):
        flag = torch.zeros(self.number_of_nodes(), dtype=torch.bool, device=self.device)
        flag[self.face[self.boundary_face_flag()].reshape(-1).long()] = True
        return flag

    def boundary_node_index(self):
        return torch.nonzero(self.boundary_node_flag())[:, 0]

    # -- geometry ---------------------------------------------------------------
    def quadrature_formula(self, q, etype='cell', qtype='legendre'):
        if etype not in ('cell', self.TD):
            raise NotImplementedError("only cell quadrature is provided")
        return simplex_rule(self.TD, q)

    def multi_index_matrix(self, p, etype, dtype=None):
        return torch.as_tensor(multi_index_matrix(p, etype), dtype=torch.int32 if dtype is None else dtype)

    def shape_function(self, bcs, p=1, *, index=_S, variables='u', mi=None):
        bc = bcs.detach().cpu().numpy() if torch.is_tensor(bcs) else np.asarray(bcs)
        return torch.as_tensor(_shape(bc, p), dtype=torch.float64, device=self.device)

    def grad_shape_function(self, bcs, p=1, *, index=_S, variables='u', mi=None):
        bc = bcs.detach().cpu().numpy() if torch.is_tensor(bcs) else np.asarray(bcs)
        R = torch.as_tensor(_grad_shape(bc, p), dtype=torch.float64, device=self.device)
        if variables == 'u':
            return R
        gl = self.grad_lambda(index=index)
        return torch.einsum('cbm, qjb -> cqjm', gl, R)

    def bc_to_point(self, bcs, index=_S):
        """(NC, NQ, GD) physical points of barycentric points bcs (NQ, TD+1),
        through fa_bc_to_point (a mesh on the host raises, as every kernel
        wrapper does)."""
        bcs = torch.as_tensor(bcs, dtype=self.ftype, device=self.device)
        cell = self.cell[index]
        from .. import _native as nat
        cell = cell.to(torch.int32).contiguous()
        nc, nv = cell.shape
        nq = bcs.shape[0]
        out = torch.empty((nc, nq, self.GD), dtype=torch.float64, device=self.device)
        nat.check(nat.lib().fa_bc_to_point(self.GD, nv, nc, nq, nat.ptr(self.node.contiguous()), nat.ptr(cell),
                                           nat.ptr(bcs.contiguous()), nat.ptr(out), nat.stream()), "fa_bc_to_point")
        return out

    def entity_barycenter(self, etype='cell', index=_S):
        e = self.entity(etype, index)
        return self.node[e.long()].mean(dim=1)

    def cell_centroid(self):
        return self.entity_barycenter('cell')

    # -- interpolation points / DOF numbering ----------------------------------
    def number_of_local_ipoints(self, p, iptype='cell'):
        d = {'cell': self.TD, 'face': self.TD - 1, 'edge': 1, 'node': 0}.get(iptype, iptype)
        return 1 if d == 0 else comb(p + d, d)

    def number_of_global_ipoints(self, p):
        if p == 1:
            return self.number_of_nodes()
        nums = [self.number_of_nodes(), self.number_of_edges()]
        if self.TD == 3:
            nums.append(self.number_of_faces())
        nums.append(self.number_of_cells())
        return _gdof(p, nums)

    def edge_to_ipoint(self, p, index=_S):
        NN, NE = self.number_of_nodes(), self.number_of_edges()
        e = self.edge
        mid = NN + (p - 1) * torch.arange(NE, dtype=e.dtype, device=self.device)[:, None] \
            + torch.arange(p - 1, dtype=e.dtype, device=self.device)[None, :]
        return torch.cat([e[:, :1], mid, e[:, 1:]], dim=1)[index]

    def interpolation_points(self, p, index=_S):
        node = self.node
        if p == 1:
            return node[index]
        GD = self.GD
        parts = [node]
        w = torch.as_tensor(multi_index_matrix(p, 1)[1:-1] / p, dtype=self.ftype, device=self.device)
        parts.append(torch.einsum('ij, ejm -> eim', w, node[self.edge.long()]).reshape(-1, GD))
        if self.TD == 3 and p > 2:
            mi = multi_index_matrix(p, 2)
            w = torch.as_tensor(mi[np.all(mi > 0, axis=1)] / p, dtype=self.ftype, device=self.device)
            parts.append(torch.einsum('ij, fjm -> fim', w, node[self.face.long()]).reshape(-1, GD))
        if p > self.TD:
            mi = multi_index_matrix(p, self.TD)
            w = torch.as_tensor(mi[np.all(mi > 0, axis=1)] / p, dtype=self.ftype, device=self.device)
            parts.append(torch.einsum('ij, cjm -> cim', w, node[self.cell.long()]).reshape(-1, GD))
        return torch.cat(parts, dim=0)[index]

    def is_boundary_dof(self, p):
        grid = self._box_grid()
        if p == 1 and grid is not None:
            # the nodes on the box's surface: those on a boundary face of the
            # full grid, from the grid position (no face numbering needed)
            idx = [torch.arange(n + 1, device=self.device) for n in grid]
            axes = torch.meshgrid(*idx, indexing='ij')
            flag = torch.zeros(axes[0].shape, dtype=torch.bool, device=self.device)
            for a, n in zip(axes, grid):
                flag |= (a == 0) | (a == n)
            return flag.reshape(-1)
        gdof = self.number_of_global_ipoints(p)
        f2d = self.face_to_ipoint(p)[self.boundary_face_index()]
        flag = torch.zeros(gdof, dtype=torch.bool, device=self.device)
        flag[f2d.reshape(-1).long()] = True
        return flag

    def _box_grid(self):
        """Cell counts per axis of a full from_box grid, else None."""
        fb = dict.get(self.meshdata, 'from_box')
        return None if fb is None else tuple(fb[2:])

    # -- L2 error (reporting; fealpy/mesh/mesh_base.py:545-608) ------------------
    def to_vtk(self, fname=None, etype='cell', index=_S):
        """Mesh (and its nodedata / celldata) as a VTK unstructured grid
        (tetrahedron_mesh.py:1536-1573, triangle_mesh.py's to_vtk).  With no
        file name, returns (node, flat (NV, v0, ...) cell list, VTK cell type,
        NC) as the reference does."""
        from .vtu import write_vtu
        ent = self.entity(etype)[index]
        nc, nv = ent.shape
        cl = torch.cat([torch.full((nc, 1), nv, dtype=ent.dtype, device=ent.device), ent], dim=1)
        ctype = {2: 3, 3: 5, 4: 10}[nv]            # VTK_LINE, VTK_TRIANGLE, VTK_TETRA
        data = {'cell': self.celldata, 'face': self.facedata, 'edge': self.edgedata}.get(etype, self.celldata)
        if fname is None:
            return self.node, cl.reshape(-1), ctype, nc
        write_vtu(fname, self.node.detach().cpu().numpy(), cl.reshape(-1).cpu().numpy(), nc, ctype,
                  nodedata=self.nodedata, celldata=data)
        return None

    def error(self, u, v, q=3, power=2, celltype=False):
        """(sum_c int_c |u - v|^power)^(1/power) (mesh_base.py:545-608).  u, v:
        cartesian or barycentric callables or (NC, NQ[, ...]) values; vector
        values (gradients: mesh.error(grad_u, uh.grad_value) is the H1
        semi-norm error) are summed over their components."""
        qf = self.quadrature_formula(q)
        bcs, ws = qf.get_quadrature_points_and_weights()
        ws = ws.to(self.device)
        ps = self.bc_to_point(bcs)
        uu = u(bcs) if getattr(u, 'coordtype', None) == 'barycentric' else (u(ps) if callable(u) else u)
        vv = v(bcs) if getattr(v, 'coordtype', None) == 'barycentric' else (v(ps) if callable(v) else v)
        uu, vv = torch.as_tensor(uu, device=self.device), torch.as_tensor(vv, device=self.device)
        if uu.ndim == 2 and vv.ndim == 3 and vv.shape[-1] == 1:
            uu = uu[..., None]
        f = torch.abs(uu - vv) ** power
        cm = self.entity_measure('cell')
        if f.ndim == 0:                             # a constant function
            e = f * cm
        elif f.ndim == 1 and f.shape[0] == self.GD:  # a constant vector
            e = cm[:, None] * f
        else:
            if f.ndim == 1:
                f = f[:, None]
            e = torch.einsum('q, cq..., c -> c...', ws.to(f.dtype), f, cm)
        if not celltype:
            return torch.sum(e) ** (1 / power)
        return torch.sum(e.reshape(e.shape[0], -1), dim=1) ** (1 / power)


class TriangleMesh(SimplexMesh):
    """fealpy/mesh/triangle_mesh.py."""
    TD = 2
    localEdge = torch.tensor([(1, 2), (2, 0), (0, 1)], dtype=torch.int64)
    localFace = localEdge

    def __init__(self, node, cell):
        super().__init__(node, cell)
        self.meshtype = 'tri'

    @classmethod
    def from_box(cls, box=(0, 1, 0, 1), nx=10, ny=10, *, threshold=None, itype=None, ftype=None, device=None):
        """Two triangles per grid square, nodes in meshgrid 'ij' order
        (triangle_mesh.py:1387-1436)."""
        device = default_device() if device is None else torch.device(device)
        itype = torch.int32 if itype is None else itype
        ftype = torch.float64 if ftype is None else ftype
        NN = (nx + 1) * (ny + 1)
        x = torch.linspace(box[0], box[1], nx + 1, dtype=ftype, device=device)
        y = torch.linspace(box[2], box[3], ny + 1, dtype=ftype, device=device)
        X, Y = torch.meshgrid(x, y, indexing='ij')
        node = torch.stack([X.reshape(-1), Y.reshape(-1)], dim=1)
        idx = torch.arange(NN, dtype=itype, device=device).reshape(nx + 1, ny + 1)
        c0 = torch.stack([idx[1:, :-1].T.reshape(-1), idx[1:, 1:].T.reshape(-1), idx[:-1, :-1].T.reshape(-1)], 1)
        c1 = torch.stack([idx[:-1, 1:].T.reshape(-1), idx[:-1, :-1].T.reshape(-1), idx[1:, 1:].T.reshape(-1)], 1)
        cell = torch.cat([c0, c1], dim=0)
        if threshold is not None:
            node, cell = _apply_threshold(node, cell, threshold)
        mesh = cls(node, cell)
        mesh.box = list(box)
        if threshold is None:           # the full structured grid (geometric multigrid reads it)
            mesh.meshdata['from_box'] = ('tri', tuple(box), nx, ny)
        return mesh

    def grad_lambda(self, index=_S, TD=2):
        """numpy_backend.py:586-598."""
        p = self.node[self.cell[index].long()]
        e0 = p[:, 2] - p[:, 1]
        e1 = p[:, 0] - p[:, 2]
        e2 = p[:, 1] - p[:, 0]
        nv = e0[:, 0] * e1[:, 1] - e0[:, 1] * e1[:, 0]
        g = torch.stack([torch.stack([-e[:, 1], e[:, 0]], -1) for e in (e0, e1, e2)], dim=1)
        return g / nv[:, None, None]

    def entity_measure(self, etype='cell', index=_S):
        node = self.node
        if etype in ('cell', 2):
            p = node[self.cell[index].long()]
            a, b = p[:, 1] - p[:, 0], p[:, 2] - p[:, 1]
            return (a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]) / 2.0
        if etype in ('edge', 'face', 1):
            e = self.edge[index].long()
            return torch.linalg.norm(node[e[:, 0]] - node[e[:, 1]], dim=-1)
        if etype in ('node', 0):
            return torch.zeros(1, dtype=self.ftype, device=self.device)
        raise ValueError(f"entity type: {etype} is wrong!")

    def face_to_ipoint(self, p, index=_S):
        return self.edge_to_ipoint(p, index)

    def cell_to_ipoint(self, p, index=_S):
        """triangle_mesh.py:218-270: edge DOFs placed in each cell with the
        orientation of the edge's first/second occurrence."""
        cell = self.cell
        if p == 1:
            return cell[index]
        dev = self.device
        mi = multi_index_matrix(p, 2)
        idx = [torch.as_tensor(np.nonzero(mi[:, k] == 0)[0], device=dev) for k in range(3)]
        f2c = self.face2cell.long()
        NN, NE, NC = self.number_of_nodes(), self.number_of_edges(), self.number_of_cells()
        e2p = self.edge_to_ipoint(p)
        ldof = comb(p + 2, 2)
        c2p = torch.zeros((NC, ldof), dtype=cell.dtype, device=dev)
        for k in range(3):                         # first occurrence: forward/reversed per local edge
            flag = f2c[:, 2] == k
            cols = idx[k] if k != 1 else torch.flip(idx[k], [0])
            c2p[f2c[flag, 0][:, None], cols] = e2p[flag]
        inner = f2c[:, 0] != f2c[:, 1]
        for k in range(3):                         # second occurrence: opposite orientation
            flag = inner & (f2c[:, 3] == k)
            cols = torch.flip(idx[k], [0]) if k != 1 else idx[k]
            c2p[f2c[flag, 1][:, None], cols] = e2p[flag]
        cdof = (p - 1) * (p - 2) // 2
        if cdof > 0:
            flag = torch.as_tensor(np.all(mi > 0, axis=1), device=dev)
            c2p[:, flag] = (NN + NE * (p - 1) + torch.arange(NC * cdof, dtype=cell.dtype, device=dev)).reshape(NC, cdof)
        return c2p[index]


class TetrahedronMesh(SimplexMesh):
    """fealpy/mesh/tetrahedron_mesh.py."""
    TD = 3
    localEdge = torch.tensor([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)], dtype=torch.int64)
    localFace = torch.tensor([(1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1)], dtype=torch.int64)
    localFace2edge = torch.tensor([(5, 4, 3), (5, 1, 2), (4, 2, 0), (3, 0, 1)], dtype=torch.int64)

    def __init__(self, node, cell):
        super().__init__(node, cell)
        self.meshtype = 'tet'

    @classmethod
    def from_box(cls, box=(0, 1, 0, 1, 0, 1), nx=10, ny=10, nz=10, threshold=None, device=None):
        """Kuhn triangulation: 6 tetrahedra per cube around the main diagonal
        (tetrahedron_mesh.py:1017-1086), nodes in meshgrid 'ij' order."""
        device = default_device() if device is None else torch.device(device)
        NN = (nx + 1) * (ny + 1) * (nz + 1)
        ax = [torch.linspace(box[2 * k], box[2 * k + 1], n + 1, dtype=torch.float64, device=device)
              for k, n in enumerate((nx, ny, nz))]
        X, Y, Z = torch.meshgrid(*ax, indexing='ij')
        node = torch.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], dim=1)
        idx = torch.arange(NN, dtype=torch.int32, device=device).reshape(nx + 1, ny + 1, nz + 1)
        nyz = (ny + 1) * (nz + 1)
        c0 = idx[:-1, :-1, :-1].reshape(-1)
        corner = [c0, c0 + nyz, c0 + nyz + nz + 1, c0 + nz + 1,
                  c0 + 1, c0 + 1 + nyz, c0 + 1 + nyz + nz + 1, c0 + 1 + nz + 1]
        cube = torch.stack(corner, dim=1)
        kuhn = torch.tensor([[0, 1, 2, 6], [0, 5, 1, 6], [0, 4, 5, 6],
                             [0, 7, 4, 6], [0, 3, 7, 6], [0, 2, 3, 6]], dtype=torch.int64, device=device)
        cell = cube[:, kuhn].reshape(-1, 4)
        if threshold is not None:
            node, cell = _apply_threshold(node, cell, threshold)
        mesh = cls(node, cell)
        mesh.box = list(box)
        if threshold is None:           # the full structured grid (geometric multigrid reads it)
            mesh.meshdata['from_box'] = ('tet', tuple(box), nx, ny, nz)
            mesh.meshdata._tagger = mesh._tag_box_faces       # the side tags on first access
        return mesh

    def _tag_box_faces(self):
        bd = self.boundary_face_index()
        n = self.face_unit_normal(bd)
        for name, (k, s) in dict(leftface=(0, -1), rightface=(0, 1), frontface=(1, -1), backface=(1, 1),
                                 bottomface=(2, -1), upface=(2, 1)).items():
            dict.__setitem__(self.meshdata, name, bd[torch.abs(n[:, k] - s) < 1e-14])

    def face_unit_normal(self, index=_S):
        f = self.face[index].long()
        v01 = self.node[f[:, 1]] - self.node[f[:, 0]]
        v02 = self.node[f[:, 2]] - self.node[f[:, 0]]
        nv = torch.linalg.cross(v01, v02, dim=1)
        return nv / torch.linalg.norm(nv, dim=1, keepdim=True)

    def face_to_edge(self, index=_S):
        f2c = self.face2cell.long()
        lfe = self.localFace2edge.to(self.device)
        return self.cell2edge[f2c[:, [0]], lfe[f2c[:, 2]]][index]

    def grad_lambda(self, index=_S, TD=3):
        """numpy_backend.py:619-629."""
        p = self.node[self.cell[index].long()]                 # (NC, 4, 3)
        a, b, c = p[:, 1] - p[:, 0], p[:, 2] - p[:, 1], p[:, 3] - p[:, 2]
        vol6 = torch.sum(a * torch.linalg.cross(b, c, dim=1), dim=1)
        out = []
        for j, k, m in self.localFace.tolist():
            vjk = p[:, k] - p[:, j]
            vjm = p[:, m] - p[:, j]
            out.append(torch.linalg.cross(vjm, vjk, dim=1) / vol6[:, None])
        return torch.stack(out, dim=1)

    def entity_measure(self, etype='cell', index=_S):
        node = self.node
        if etype in ('cell', 3):
            c = self.cell[index].long()
            v01 = node[c[:, 1]] - node[c[:, 0]]
            v02 = node[c[:, 2]] - node[c[:, 0]]
            v03 = node[c[:, 3]] - node[c[:, 0]]
            return torch.sum(v03 * torch.linalg.cross(v01, v02, dim=1), dim=1) / 6.0
        if etype in ('face', 2):
            f = self.face[index].long()
            nv = torch.linalg.cross(node[f[:, 1]] - node[f[:, 0]], node[f[:, 2]] - node[f[:, 0]], dim=1)
            return torch.sqrt(torch.sum(nv * nv, dim=1)) / 2.0
        if etype in ('edge', 1):
            e = self.edge[index].long()
            return torch.linalg.norm(node[e[:, 0]] - node[e[:, 1]], dim=-1)
        if etype in ('node', 0):
            return torch.zeros(1, dtype=self.ftype, device=self.device)
        raise ValueError(f"entity type: {etype} is wrong!")

    def face_to_ipoint(self, p, index=_S):
        """tetrahedron_mesh.py:348-386: edge DOFs of each face in the face's
        orientation, then face-interior DOFs."""
        dev = self.device
        fdof = comb(p + 2, 2)
        NN, NE, NF = self.number_of_nodes(), self.number_of_edges(), self.number_of_faces()
        face, edge = self.face.long(), self.edge.long()
        f2e = self.face_to_edge().long()
        e2p = self.edge_to_ipoint(p)
        f2p = torch.zeros((NF, fdof), dtype=self.itype, device=dev)
        mi = multi_index_matrix(p, 2)
        on_edge = mi == 0
        fwd = torch.arange(p + 1, device=dev)
        rev = torch.flip(fwd, [0])
        first = (1, 0, 0)
        for i in range(3):
            same = face[:, first[i]] == edge[f2e[:, i], 0]
            order = torch.where(same[:, None], fwd[None, :], rev[None, :])
            cols = torch.as_tensor(np.nonzero(on_edge[:, i])[0], device=dev)
            f2p[:, cols] = torch.gather(e2p[f2e[:, i]], 1, order)
        if p > 2:
            fidof = fdof - 3 * p
            inner = torch.as_tensor(~on_edge.any(axis=1), device=dev)
            f2p[:, inner] = (NN + (p - 1) * NE
                             + torch.arange(NF * fidof, dtype=self.itype, device=dev)).reshape(NF, fidof)
        return f2p[index]

    def cell_to_ipoint(self, p, index=_S):
        """tetrahedron_mesh.py:388-441: face DOFs placed through the permutation
        between the cell's local face and the stored face, then cell-interior DOFs."""
        cell = self.cell
        if p == 1:
            return cell[index]
        dev = self.device
        NN, NE, NF, NC = self.number_of_nodes(), self.number_of_edges(), self.number_of_faces(), self.number_of_cells()
        fdof = comb(p + 2, 2)
        ldof = comb(p + 3, 3)
        c2f = self.cell2face.long()
        c2p = torch.zeros((NC, ldof), dtype=self.itype, device=dev)
        f2p = self.face_to_ipoint(p)
        m2 = torch.as_tensor(multi_index_matrix(p, 2).T.copy(), dtype=torch.int64, device=dev)   # (3, fdof)
        m3 = multi_index_matrix(p, 3).T                                                          # (4, ldof)
        on_face = m3 == 0
        frank = torch.argsort(torch.argsort(self.face.long(), dim=1), dim=1)
        ar = torch.arange(NC, device=dev)[:, None]
        for i in range(4):
            others = [k for k in range(4) if k != i]
            idxj = torch.argsort(cell[:, others].long(), dim=1)
            idxi = frank[c2f[:, i]]
            order = idxj[ar, idxi]                       # (NC, 3)
            mi = m2[order]                               # (NC, 3, fdof)
            k = mi[:, 1] + mi[:, 2]
            a = k * (k + 1) // 2 + mi[:, 2]
            cols = torch.as_tensor(np.nonzero(on_face[i])[0], device=dev)
            c2p[:, cols] = torch.gather(f2p[c2f[:, i]], 1, a).to(self.itype)
        if p > 3:
            idof = ldof - 4 - 6 * (p - 1) - 4 * (fdof - 3 * p)
            inner = torch.as_tensor(~on_face.any(axis=0), device=dev)
            base = NN + (p - 1) * NE + (fdof - 3 * p) * NF
            c2p[:, inner] = (base + torch.arange(NC * idof, dtype=self.itype, device=dev)).reshape(NC, idof)
        return c2p[index]


def _apply_threshold(node, cell, threshold):
    NN = node.shape[0]
    bc = node[cell.long()].sum(dim=1) / cell.shape[1]
    cell = cell[~threshold(bc)]
    valid = torch.zeros(NN, dtype=torch.bool, device=node.device)
    valid[cell.reshape(-1).long()] = True
    node = node[valid]
    idxmap = torch.zeros(NN, dtype=cell.dtype, device=node.device)
    idxmap[valid] = torch.arange(int(valid.sum()), dtype=cell.dtype, device=node.device)
    return node, idxmap[cell.long()]
