"""Minimal meshio-compatible mesh container and PLY / STL readers.

The reference takes `meshio.Mesh` objects (points, cells, point_data) on its
public entry points; meshio is not a dependency here, so this module provides an
object with the same attributes the hot path reads (`points`, `cells[i].type`,
`cells[i].data`, `point_data`, `cell_data`) and readers for the PLY / STL files
the reference ships as test data.  A real `meshio.Mesh` works too wherever a
mesh is accepted.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np


@dataclass
class CellBlock:
    type: str
    data: np.ndarray

    def __post_init__(self):
        self.data = np.asarray(self.data)

    def __len__(self):
        return len(self.data)


@dataclass
class Mesh:
    points: np.ndarray
    cells: list = field(default_factory=list)
    point_data: dict = field(default_factory=dict)
    cell_data: dict = field(default_factory=dict)

    def __post_init__(self):
        self.points = np.asarray(self.points)
        blocks = []
        for c in self.cells:
            if isinstance(c, CellBlock) or hasattr(c, "type"):
                blocks.append(c)
            else:
                blocks.append(CellBlock(c[0], np.asarray(c[1])))
        self.cells = blocks


def triangle_cells(mesh) -> np.ndarray:
    """The (T,3) vertex indices of the mesh's last triangle cell block
    (convertTriangles keeps the last one it finds, rayfunctions.py:537-540)."""
    tri = None
    for c in mesh.cells:
        if c.type == "triangle":
            tri = np.asarray(c.data)
    if tri is None:
        return np.empty((0, 3), np.int64)
    return tri.reshape(-1, 3)


_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1",
    "short": "i2", "int16": "i2", "ushort": "u2", "uint16": "u2",
    "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}


def read_ply(path: str) -> Mesh:
    """Read an ASCII or binary PLY file with vertex x/y/z and triangle faces."""
    with open(path, "rb") as f:
        data = f.read()
    end = data.index(b"end_header")
    end = data.index(b"\n", end) + 1
    header = data[:end].decode("ascii", errors="replace").splitlines()
    fmt = None
    elements = []
    for line in header:
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            elements.append([tok[1], int(tok[2]), []])
        elif tok[0] == "property":
            if tok[1] == "list":
                elements[-1][2].append((tok[4], "list", _PLY_TYPES[tok[2]], _PLY_TYPES[tok[3]]))
            else:
                elements[-1][2].append((tok[2], _PLY_TYPES[tok[1]], None, None))
    body = data[end:]
    points = np.empty((0, 3))
    faces = []
    if fmt == "ascii":
        lines = body.decode("ascii").split("\n")
        li = 0
        for name, count, props in elements:
            rows = []
            for _ in range(count):
                vals = lines[li].split()
                li += 1
                rows.append(vals)
            if name == "vertex":
                names = [p[0] for p in props]
                arr = np.array([[float(v) for v in r[: len(names)]] for r in rows])
                points = arr[:, [names.index("x"), names.index("y"), names.index("z")]]
            elif name == "face":
                for r in rows:
                    n = int(r[0])
                    faces.append([int(v) for v in r[1 : 1 + n]])
    else:
        endian = "<" if fmt == "binary_little_endian" else ">"
        off = 0
        for name, count, props in elements:
            if all(p[1] != "list" for p in props):
                dt = np.dtype([(p[0], endian + p[1]) for p in props])
                arr = np.frombuffer(body, dtype=dt, count=count, offset=off)
                off += dt.itemsize * count
                if name == "vertex":
                    points = np.stack([arr["x"], arr["y"], arr["z"]], axis=1).astype(np.float64)
            else:
                for _ in range(count):
                    row = []
                    for pname, kind, ctype, itype in props:
                        if kind == "list":
                            cs = np.dtype(ctype).itemsize
                            n = int(np.frombuffer(body, endian + ctype, 1, off)[0])
                            off += cs
                            isz = np.dtype(itype).itemsize
                            row = np.frombuffer(body, endian + itype, n, off).astype(np.int64)
                            off += isz * n
                        else:
                            off += np.dtype(kind).itemsize
                    if name == "face":
                        faces.append(list(row))
    tris = []
    for f in faces:
        for i in range(1, len(f) - 1):
            tris.append([f[0], f[i], f[i + 1]])
    cells = [CellBlock("triangle", np.asarray(tris, dtype=np.int64).reshape(-1, 3))]
    return Mesh(points=np.asarray(points, np.float64), cells=cells)


def read_stl(path: str) -> Mesh:
    """Read a binary or ASCII STL file (vertices are not merged)."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:5] == b"solid" and b"facet" in data[:512]:
        verts = []
        for line in data.decode("ascii", errors="replace").splitlines():
            tok = line.split()
            if tok and tok[0] == "vertex":
                verts.append([float(v) for v in tok[1:4]])
        pts = np.asarray(verts, np.float64).reshape(-1, 3)
    else:
        n = struct.unpack_from("<I", data, 80)[0]
        dt = np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")])
        rec = np.frombuffer(data, dtype=dt, count=n, offset=84)
        pts = rec["v"].reshape(-1, 3).astype(np.float64)
    tris = np.arange(len(pts), dtype=np.int64).reshape(-1, 3)
    return Mesh(points=pts, cells=[CellBlock("triangle", tris)])


def read(path: str) -> Mesh:
    p = path.lower()
    if p.endswith(".ply"):
        return read_ply(path)
    if p.endswith(".stl"):
        return read_stl(path)
    raise ValueError(f"unsupported mesh file: {path}")
