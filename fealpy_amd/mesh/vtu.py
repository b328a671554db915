"""VTK XML unstructured-grid (.vtu) writer for simplex meshes.

The reference writes through the vtk package (fealpy/mesh/vtk_extent.py:20-80,
called by to_vtk, fealpy/mesh/tetrahedron_mesh.py:1536-1573); that package is
not a dependency here, so the file is written directly: points, the
connectivity of the (NV, v0, ..., v_{NV-1}) cell list the reference builds,
offsets, VTK cell types, then point and cell data arrays, as base64-encoded
little-endian binary blocks with a UInt64 byte-count header (the VTK XML
"binary" format that ParaView and meshio read)."""
import base64
import struct

import numpy as np

_VTK_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32", np.dtype(np.int64): "Int64",
              np.dtype(np.int32): "Int32", np.dtype(np.uint8): "UInt8", np.dtype(np.int8): "Int8"}


def _block(a):
    a = np.ascontiguousarray(a)
    raw = a.tobytes()
    return base64.b64encode(struct.pack("<Q", len(raw)) + raw).decode("ascii")


def _array_xml(name, a, indent="        "):
    a = np.asarray(a)
    if a.dtype == np.bool_:
        a = a.astype(np.uint8)
    if a.dtype not in _VTK_TYPES:
        a = a.astype(np.float64)
    a = a.astype(a.dtype.newbyteorder("<"), copy=False)
    ncomp = 1 if a.ndim == 1 else int(np.prod(a.shape[1:]))
    return (f'{indent}<DataArray type="{_VTK_TYPES[np.dtype(a.dtype.name)]}" Name="{name}" '
            f'NumberOfComponents="{ncomp}" format="binary">{_block(a.reshape(a.shape[0], -1) if a.ndim > 1 else a)}'
            f'</DataArray>\n')


def write_vtu(fname, node, cell_list, nc, cell_type, nodedata=None, celldata=None):
    """node (NN, GD); cell_list the flat (NV, v0, ..., v_{NV-1}) records of
    nc cells of VTK type cell_type; nodedata / celldata: name -> array
    whose first axis runs over the nodes / cells."""
    node = np.asarray(node, dtype=np.float64)
    if node.shape[1] == 2:                      # VTK points are 3-D
        node = np.concatenate([node, np.zeros((node.shape[0], 1))], axis=1)
    cl = np.asarray(cell_list, dtype=np.int64).reshape(nc, -1)
    nv = cl[:, 0]
    if not np.all(nv == nv[0]):
        raise ValueError("write_vtu takes cells with one vertex count")
    conn = cl[:, 1:].reshape(-1)
    offsets = np.cumsum(nv)
    types = np.full(nc, cell_type, dtype=np.uint8)
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n',
           '  <UnstructuredGrid>\n',
           f'    <Piece NumberOfPoints="{node.shape[0]}" NumberOfCells="{nc}">\n',
           '      <Points>\n', _array_xml("Points", node), '      </Points>\n',
           '      <Cells>\n', _array_xml("connectivity", conn), _array_xml("offsets", offsets),
           _array_xml("types", types), '      </Cells>\n']
    for tag, data in (("PointData", nodedata), ("CellData", celldata)):
        out.append(f'      <{tag}>\n')
        for key, val in (data or {}).items():
            if val is None:
                continue
            v = val.detach().cpu().numpy() if hasattr(val, "detach") else np.asarray(val)
            if v.ndim == 2 and v.shape[1] == 2:
                v = np.concatenate([v, np.zeros((v.shape[0], 1), dtype=v.dtype)], axis=1)
            out.append(_array_xml(key, v))
        out.append(f'      </{tag}>\n')
    out += ['    </Piece>\n', '  </UnstructuredGrid>\n', '</VTKFile>\n']
    with open(fname, "w") as fh:
        fh.writelines(out)


def read_vtu_arrays(fname):
    """The data arrays of a .vtu written by write_vtu: {name: ndarray}
    (for checks; not a general VTK reader)."""
    import xml.etree.ElementTree as ET
    root = ET.parse(fname).getroot()
    rev = {v: k for k, v in _VTK_TYPES.items()}
    out = {}
    for da in root.iter("DataArray"):
        raw = base64.b64decode(da.text)
        n = struct.unpack("<Q", raw[:8])[0]
        a = np.frombuffer(raw[8:8 + n], dtype=rev[da.get("type")].newbyteorder("<"))
        k = int(da.get("NumberOfComponents", "1"))
        out[da.get("Name")] = a.reshape(-1, k) if k > 1 else a
    piece = root.find(".//Piece")
    out["_counts"] = (int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells")))
    return out
