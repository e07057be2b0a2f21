"""Hexahedral VTU output of cell fields (src/gridfoam/io/vtu.py:7-134).

Every leaf cell becomes one VTK_HEXAHEDRON with eight points of its own (no
point sharing), in VTK's node order, exactly as the reference builds its
pyvista UnstructuredGrid; fields are written as cell data.  The file is VTK
XML (UnstructuredGrid, little-endian, UInt64 headers, inline base64 binary),
written with numpy alone; ``read_vtu`` reads it back.
"""

from __future__ import annotations

import base64
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import Dict, Iterable, Optional

import numpy as np
import torch

VTK_HEXAHEDRON = 12

# (sx, sy, sz) signs of the eight hexahedron nodes in VTK order (vtu.py:33-41)
_NODE_SIGNS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                        [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)

_VTK_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32", np.dtype(np.int64): "Int64",
              np.dtype(np.int32): "Int32", np.dtype(np.uint8): "UInt8"}
_NP_TYPES = {v: k for k, v in _VTK_TYPES.items()}


@dataclass
class HexMesh:
    """Unstructured hexahedral grid: points [8C, 3], connectivity [8C],
    offsets [C], types [C] and cell data name -> [C] or [C, k]."""

    points: np.ndarray
    connectivity: np.ndarray
    offsets: np.ndarray
    types: np.ndarray
    cell_data: Dict[str, np.ndarray] = field(default_factory=dict)

    @property
    def n_cells(self) -> int:
        return int(self.types.shape[0])


def to_unstructured_grid(grid) -> HexMesh:
    """vtu.py:7-77: corner points from cell centres and sizes (formed on the
    grid's device, then copied to the host)."""
    c = grid.cell_centers.to(torch.float64)
    h = grid.cell_sizes.to(torch.float64) / 2.0
    signs = torch.as_tensor(_NODE_SIGNS, device=c.device)
    pts = c[:, None, :] + signs[None, :, :] * h[:, None, :]
    C = int(c.shape[0])
    points = pts.reshape(-1, 3).cpu().numpy()
    conn = np.arange(8 * C, dtype=np.int64)
    offsets = np.arange(8, 8 * C + 1, 8, dtype=np.int64)
    types = np.full(C, VTK_HEXAHEDRON, dtype=np.uint8)
    return HexMesh(points, conn, offsets, types)


def update_export_cell_data(ugrid: HexMesh, grid, fields: Iterable = ()) -> None:
    """vtu.py:80-104: cell data from the given fields whose ``export`` flag
    (the condition's ``export``) is set."""
    for f in fields:
        if not getattr(f, "export", True):
            continue
        values = f.data.detach().cpu().numpy()
        if values.ndim > 1:
            values = values.reshape(grid.num_cells, -1)
        ugrid.cell_data[f.name] = values


def save_export_fields_as_vtu(grid, output_path: str, *, ugrid: Optional[HexMesh] = None,
                              fields: Iterable = ()) -> HexMesh:
    """vtu.py:107-134: build (or reuse) the hexahedral grid, attach the
    exported fields and write the file."""
    target = ugrid if ugrid is not None else to_unstructured_grid(grid)
    update_export_cell_data(target, grid, fields)
    write_vtu(target, output_path)
    return target


def _encode(a: np.ndarray) -> str:
    raw = np.ascontiguousarray(a).tobytes()
    return base64.b64encode(np.uint64(len(raw)).tobytes() + raw).decode("ascii")


def _array_xml(a: np.ndarray, name: Optional[str] = None) -> str:
    comps = 1 if a.ndim == 1 else int(a.shape[1])
    attrs = f'type="{_VTK_TYPES[a.dtype]}"'
    if name is not None:
        attrs += f' Name="{name}"'
    if comps > 1:
        attrs += f' NumberOfComponents="{comps}"'
    return f'<DataArray {attrs} format="binary">{_encode(a)}</DataArray>\n'


def write_vtu(mesh: HexMesh, path: str) -> None:
    """VTK XML UnstructuredGrid, inline base64 binary with UInt64 byte counts."""
    parts = ['<?xml version="1.0"?>\n',
             '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n',
             '<UnstructuredGrid>\n',
             f'<Piece NumberOfPoints="{mesh.points.shape[0]}" NumberOfCells="{mesh.n_cells}">\n',
             '<Points>\n', _array_xml(mesh.points.astype(np.float64)), '</Points>\n',
             '<Cells>\n', _array_xml(mesh.connectivity.astype(np.int64), "connectivity"),
             _array_xml(mesh.offsets.astype(np.int64), "offsets"), _array_xml(mesh.types.astype(np.uint8), "types"),
             '</Cells>\n', '<CellData>\n']
    for name, values in mesh.cell_data.items():
        parts.append(_array_xml(np.asarray(values), name))
    parts += ['</CellData>\n', '</Piece>\n', '</UnstructuredGrid>\n', '</VTKFile>\n']
    with open(path, "w") as f:
        f.writelines(parts)


def _decode(el: ET.Element) -> np.ndarray:
    raw = base64.b64decode(el.text.strip())
    n = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
    a = np.frombuffer(raw[8:8 + n], dtype=_NP_TYPES[el.get("type")]).copy()
    comps = int(el.get("NumberOfComponents", "1"))
    return a.reshape(-1, comps) if comps > 1 else a


def read_vtu(path: str) -> HexMesh:
    """Read a file written by ``write_vtu``."""
    piece = ET.parse(path).getroot().find("UnstructuredGrid").find("Piece")
    points = _decode(piece.find("Points").find("DataArray"))
    cells = {el.get("Name"): _decode(el) for el in piece.find("Cells").findall("DataArray")}
    data = {el.get("Name"): _decode(el) for el in piece.find("CellData").findall("DataArray")}
    return HexMesh(points, cells["connectivity"], cells["offsets"], cells["types"], data)
