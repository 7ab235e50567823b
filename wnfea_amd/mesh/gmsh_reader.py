"""Reader of gmsh ``.msh`` files with 10-node tetrahedra (gmsh element type 11).

The reference meshes STEP/CAD geometry through the gmsh Python API and keeps
the type-11 tetrahedra, mapping node tags to 0-based indices in the order
``getNodes()`` returns them and swapping the last two mid-edge nodes
(``wnfea/mesh/gmsh_mesher.py:221-256``): gmsh numbers the edges of a
10-node tetrahedron (0-1, 1-2, 2-0, 3-0, 2-3, 1-3), Abaqus C3D10 and
VTK_QUADRATIC_TETRA (0-1, 1-2, 2-0, 0-3, 1-3, 2-3).  gmsh itself is not
available here, so this module reads the files directly: MSH 4.1 and 2.2,
ASCII and binary.  Nodes keep their order of appearance in ``$Nodes`` (the
order of ``getNodes()``), elements the order of ``$Elements``; other element
types (points, lines, surface triangles) are skipped.  ``load_msh`` fills an
``FEAModel`` as the reference's ``_populate_model`` does
(``gmsh_mesher.py:258-285``).
"""
from __future__ import annotations

import struct

import numpy as np

from ..model import FEAModel
from ..properties.materials import MaterialDef

GMSH_TET10 = 11
# column j of the C3D10 connectivity is column GMSH_TO_C3D10[j] of the gmsh one
GMSH_TO_C3D10 = np.array([0, 1, 2, 3, 4, 5, 6, 7, 9, 8])
# nodes per element for the gmsh types a file may carry next to the tets
_NODES_PER_TYPE = {1: 2, 2: 3, 3: 4, 4: 4, 5: 8, 6: 6, 7: 5, 8: 3, 9: 6, 10: 9, 11: 10, 12: 27, 13: 18, 14: 14,
                   15: 1, 16: 8, 17: 20, 18: 15, 19: 13, 20: 9, 21: 10, 22: 12, 23: 15, 24: 15, 25: 21, 26: 4,
                   27: 5, 28: 6, 29: 20, 30: 35, 31: 56}


class GmshReadError(ValueError):
    """The file is not a gmsh mesh this reader understands, or holds no type-11 tetrahedra."""


def _sections(data: bytes) -> dict:
    """Byte ranges of the $Section ... $EndSection blocks (first occurrence)."""
    out = {}
    pos = 0
    while True:
        i = data.find(b"$", pos)
        if i < 0:
            break
        j = data.find(b"\n", i)
        if j < 0:
            break
        name = data[i + 1:j].strip().decode("ascii", "replace")
        if name.startswith("End") or not name:
            pos = j + 1
            continue
        end = data.find(b"$End" + name.encode(), j)
        if end < 0:
            raise GmshReadError(f"section ${name} is not closed")
        out.setdefault(name, (j + 1, end))
        pos = end + 4 + len(name)
    return out


class _Bin:
    """Sequential reader over the binary payload of a section."""

    def __init__(self, data: bytes, pos: int, endian: str, size_t: int):
        self.d, self.p, self.e = data, pos, endian
        self.st = "Q" if size_t == 8 else "I"

    def ints(self, n: int, fmt: str = "i") -> np.ndarray:
        dt = np.dtype(self.e + {"i": "i4", "I": "u4", "Q": "u8", "q": "i8"}[fmt])
        a = np.frombuffer(self.d, dtype=dt, count=n, offset=self.p)
        self.p += n * dt.itemsize
        return a

    def size_t(self, n: int) -> np.ndarray:
        return self.ints(n, self.st)

    def doubles(self, n: int) -> np.ndarray:
        a = np.frombuffer(self.d, dtype=np.dtype(self.e + "f8"), count=n, offset=self.p)
        self.p += 8 * n
        return a


def _read_v4(data: bytes, sec: dict, binary: bool, endian: str, size_t: int):
    a, b = sec["Nodes"]
    tags_l, xyz_l = [], []
    if binary:
        r = _Bin(data, a, endian, size_t)
        n_blocks, _n, _lo, _hi = (int(v) for v in r.size_t(4))
        for _ in range(n_blocks):
            _dim, _tag, parametric = (int(v) for v in r.ints(3))
            cnt = int(r.size_t(1)[0])
            tags_l.append(r.size_t(cnt).astype(np.int64))
            dim = _dim
            ncoord = 3 + (dim if parametric else 0)
            xyz_l.append(r.doubles(cnt * ncoord).reshape(cnt, ncoord)[:, :3])
    else:
        tok = data[a:b].split()
        n_blocks = int(tok[0])
        k = 4
        for _ in range(n_blocks):
            dim, parametric, cnt = int(tok[k]), int(tok[k + 2]), int(tok[k + 3])
            k += 4
            tags_l.append(np.array(tok[k:k + cnt], dtype=np.int64))
            k += cnt
            ncoord = 3 + (dim if parametric else 0)
            xyz_l.append(np.array(tok[k:k + cnt * ncoord], dtype=np.float64).reshape(cnt, ncoord)[:, :3])
            k += cnt * ncoord
    tags = np.concatenate(tags_l) if tags_l else np.zeros(0, np.int64)
    coords = np.concatenate(xyz_l) if xyz_l else np.zeros((0, 3))
    a, b = sec["Elements"]
    conn_l = []
    if binary:
        r = _Bin(data, a, endian, size_t)
        n_blocks = int(r.size_t(4)[0])
        for _ in range(n_blocks):
            _dim, _tag, etype = (int(v) for v in r.ints(3))
            cnt = int(r.size_t(1)[0])
            npe = _NODES_PER_TYPE.get(etype)
            if npe is None:
                raise GmshReadError(f"unknown gmsh element type {etype}")
            blk = r.size_t(cnt * (npe + 1)).reshape(cnt, npe + 1).astype(np.int64)
            if etype == GMSH_TET10:
                conn_l.append(blk[:, 1:])
    else:
        tok = data[a:b].split()
        n_blocks = int(tok[0])
        k = 4
        for _ in range(n_blocks):
            etype, cnt = int(tok[k + 2]), int(tok[k + 3])
            k += 4
            npe = _NODES_PER_TYPE.get(etype)
            if npe is None:
                raise GmshReadError(f"unknown gmsh element type {etype}")
            if etype == GMSH_TET10:
                conn_l.append(np.array(tok[k:k + cnt * (npe + 1)], dtype=np.int64).reshape(cnt, npe + 1)[:, 1:])
            k += cnt * (npe + 1)
    return tags, coords, conn_l


def _read_v2(data: bytes, sec: dict, binary: bool, endian: str):
    a, b = sec["Nodes"]
    if binary:
        j = data.find(b"\n", a)
        n = int(data[a:j])
        rec = np.dtype([("tag", endian + "i4"), ("xyz", endian + "f8", 3)])
        arr = np.frombuffer(data, dtype=rec, count=n, offset=j + 1)
        tags, coords = arr["tag"].astype(np.int64), arr["xyz"].astype(np.float64)
    else:
        tok = data[a:b].split()
        n = int(tok[0])
        arr = np.array(tok[1:1 + 4 * n], dtype=np.float64).reshape(n, 4)
        tags, coords = arr[:, 0].astype(np.int64), arr[:, 1:].copy()
    a, b = sec["Elements"]
    conn_l = []
    if binary:
        j = data.find(b"\n", a)
        total = int(data[a:j])
        p = j + 1
        done = 0
        while done < total:
            etype, cnt, ntags = struct.unpack_from(endian + "3i", data, p)
            p += 12
            npe = _NODES_PER_TYPE.get(etype)
            if npe is None:
                raise GmshReadError(f"unknown gmsh element type {etype}")
            width = 1 + ntags + npe
            blk = np.frombuffer(data, dtype=np.dtype(endian + "i4"), count=cnt * width, offset=p).reshape(cnt, width)
            p += 4 * cnt * width
            if etype == GMSH_TET10:
                conn_l.append(blk[:, 1 + ntags:].astype(np.int64))
            done += cnt
    else:
        lines = data[a:b].decode("ascii").split("\n")
        n = int(lines[0])
        rows = []
        for ln in lines[1:1 + n]:
            v = ln.split()
            if int(v[1]) == GMSH_TET10:
                rows.append([int(t) for t in v[3 + int(v[2]):]])
        if rows:
            conn_l.append(np.array(rows, dtype=np.int64))
    return tags, coords, conn_l


def read_msh(path) -> tuple[np.ndarray, np.ndarray]:
    """(coords (N, 3) float64, elements (M, 10) int64 in C3D10 order) of a
    gmsh .msh file (MSH 4.1 or 2.2, ASCII or binary)."""
    with open(path, "rb") as f:
        data = f.read()
    sec = _sections(data)
    if "MeshFormat" not in sec or "Nodes" not in sec or "Elements" not in sec:
        raise GmshReadError(f"{path}: not a gmsh mesh ($MeshFormat/$Nodes/$Elements missing)")
    a, b = sec["MeshFormat"]
    head = data[a:b].split(b"\n")[0].split()
    version, binary, size_t = float(head[0]), int(head[1]) == 1, int(head[2])
    endian = "<"
    if binary:
        one = data[a + data[a:b].find(b"\n") + 1:][:4]
        endian = "<" if struct.unpack("<i", one)[0] == 1 else ">"
    if version >= 4.0:
        tags, coords, conn_l = _read_v4(data, sec, binary, endian, size_t)
    elif version >= 2.0:
        tags, coords, conn_l = _read_v2(data, sec, binary, endian)
    else:
        raise GmshReadError(f"{path}: MSH version {version} is not supported")
    if not conn_l:
        raise GmshReadError(f"{path}: no 10-node tetrahedra (gmsh type 11) in the file")
    raw = np.vstack(conn_l)
    # tags -> 0-based index in order of appearance
    order = np.argsort(tags, kind="stable")
    st = tags[order]
    if len(st) > 1 and np.any(st[1:] == st[:-1]):
        raise GmshReadError(f"{path}: duplicate node tags")
    pos = np.searchsorted(st, raw)
    if np.any(pos >= len(st)) or np.any(st[np.minimum(pos, len(st) - 1)] != raw):
        raise GmshReadError(f"{path}: an element references a node tag absent from $Nodes")
    conn = order[pos]
    return np.ascontiguousarray(coords, dtype=np.float64), np.ascontiguousarray(conn[:, GMSH_TO_C3D10])


def load_msh(path, model: FEAModel | None = None, material_name: str = "Steel"):
    """Read a .msh file into an FEAModel (``gmsh_mesher.py:258-285``): the
    mesh replaces the model's, every element gets ``material_name``, and a
    structural steel (210 GPa, 0.3, 250 MPa) is added under that name when
    the model does not define it.  Returns (nodes, elements, model)."""
    nodes, elements = read_msh(path)
    if model is None:
        model = FEAModel()
    model.clear_mesh()
    model.mesh_nodes = nodes
    model.solid_elements = elements
    model.source_file = str(path)
    model.solid_materials.update(dict.fromkeys(range(len(elements)), material_name))
    if material_name not in model.materials:
        model.materials[material_name] = MaterialDef(name=material_name, youngs_modulus=210.0e9,
                                                     poissons_ratio=0.3, yield_strength=250.0e6)
    return nodes, elements, model


def get_nodes_in_box(nodes, xmin=-np.inf, xmax=np.inf, ymin=-np.inf, ymax=np.inf, zmin=-np.inf, zmax=np.inf,
                     tol: float = 1e-5) -> np.ndarray:
    """Indices of nodes inside or on an axis-aligned box (``gmsh_mesher.py:287-305``)."""
    p = np.asarray(nodes)
    m = ((p[:, 0] >= xmin - tol) & (p[:, 0] <= xmax + tol) & (p[:, 1] >= ymin - tol) & (p[:, 1] <= ymax + tol)
         & (p[:, 2] >= zmin - tol) & (p[:, 2] <= zmax + tol))
    return np.where(m)[0]


def get_nodes_on_plane(nodes, axis: str = "x", coord: float = 0.0, tol: float = 1e-5) -> np.ndarray:
    """Indices of nodes on a coordinate plane (``gmsh_mesher.py:307-323``)."""
    col = {"x": 0, "y": 1, "z": 2}[axis.lower()]
    return np.where(np.abs(np.asarray(nodes)[:, col] - coord) <= tol)[0]
