"""Mesh ingestion (gmsh .msh, type-11 tetrahedra) and VTU export, on the CPU.

The .msh files are written here from a generated C3D10 box in each of the four
encodings (MSH 4.1 / 2.2, ASCII / binary), with shuffled non-contiguous node
tags and extra surface/point elements, the mid-edge slots placed by the gmsh
edge definition (4: 0-1, 5: 1-2, 6: 2-0, 7: 3-0, 8: 2-3, 9: 3-1) found
geometrically, not through a permutation table.  The reader's output is then
checked against the Abaqus C3D10 edge definition, again geometrically, so a
missing or wrong 8<->9 swap fails.  VTU files are parsed back with an XML
parser and checked array by array."""
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

GMSH_EDGES = [(0, 1), (1, 2), (2, 0), (3, 0), (2, 3), (3, 1)]      # slots 4..9 of gmsh type 11
C3D10_EDGES = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]     # slots 4..9 of Abaqus C3D10


def _box():
    from wnfea_amd.mesh.solid_generator import generate_c3d10_structured_block
    m, _ = generate_c3d10_structured_block(0.4, 0.1, 0.1, 4, 2, 2, tip_load_total=-1000.0, tets_per_hex=6,
                                           load_direction="z")
    return np.asarray(m.mesh_nodes, dtype=np.float64), np.asarray(m.solid_elements, dtype=np.int64)


def _to_gmsh(nodes, elems):
    """Gmsh-ordered connectivity: each mid-edge slot takes the element node at
    the midpoint of that gmsh edge."""
    out = np.empty_like(elems)
    out[:, :4] = elems[:, :4]
    for e, el in enumerate(elems):
        mids = el[4:]
        for s, (a, b) in enumerate(GMSH_EDGES):
            target = 0.5 * (nodes[el[a]] + nodes[el[b]])
            k = int(np.argmin(np.linalg.norm(nodes[mids] - target, axis=1)))
            assert np.allclose(nodes[mids[k]], target)
            out[e, 4 + s] = mids[k]
    return out


def _write(path, nodes, gconn, version, binary, seed=0):
    rng = np.random.default_rng(seed)
    n = len(nodes)
    tags = rng.choice(np.arange(1, 10 * n), size=n, replace=False)   # non-contiguous, shuffled
    order = rng.permutation(n)                                       # file order of the nodes
    t_conn = tags[gconn]
    tris = t_conn[:3, [0, 1, 2, 4, 5, 6]]                              # a few 6-node triangles (type 9)
    pts = tags[:2].reshape(-1, 1)                                       # and points (type 15)
    blocks = [(15, pts), (9, tris), (11, t_conn[: len(t_conn) // 2]), (11, t_conn[len(t_conn) // 2:])]
    with open(path, "wb") as f:
        w = f.write
        if version == 4:
            w(b"$MeshFormat\n4.1 %d 8\n" % int(binary))
            if binary:
                w(struct.pack("<i", 1) + b"\n")
            w(b"$EndMeshFormat\n$Nodes\n")
            halves = [order[: n // 2], order[n // 2:]]
            if binary:
                w(struct.pack("<4Q", 2, n, int(tags.min()), int(tags.max())))
                for dim, h in zip((2, 3), halves):
                    w(struct.pack("<3i", dim, 1, 0) + struct.pack("<Q", len(h)))
                    w(tags[h].astype("<u8").tobytes())
                    w(nodes[h].astype("<f8").tobytes())
                w(b"\n$EndNodes\n$Elements\n")
                ne = sum(len(b[1]) for b in blocks)
                w(struct.pack("<4Q", len(blocks), ne, 1, ne))
                et = 1
                for etype, c in blocks:
                    w(struct.pack("<3i", 3 if etype == 11 else 0, 1, etype) + struct.pack("<Q", len(c)))
                    rows = np.column_stack([np.arange(et, et + len(c)), c]).astype("<u8")
                    w(rows.tobytes())
                    et += len(c)
                w(b"\n$EndElements\n")
            else:
                w(b"2 %d %d %d\n" % (n, tags.min(), tags.max()))
                for dim, h in zip((2, 3), halves):
                    w(b"%d 1 0 %d\n" % (dim, len(h)))
                    w("".join(f"{t}\n" for t in tags[h]).encode())
                    w("".join(f"{x!r} {y!r} {z!r}\n" for x, y, z in nodes[h].tolist()).encode())
                w(b"$EndNodes\n$Elements\n")
                ne = sum(len(b[1]) for b in blocks)
                w(b"%d %d 1 %d\n" % (len(blocks), ne, ne))
                et = 1
                for etype, c in blocks:
                    w(b"%d 1 %d %d\n" % (3 if etype == 11 else 0, etype, len(c)))
                    for row in c:
                        w((f"{et} " + " ".join(map(str, row)) + "\n").encode())
                        et += 1
                w(b"$EndElements\n")
        else:
            w(b"$MeshFormat\n2.2 %d 8\n" % int(binary))
            if binary:
                w(struct.pack("<i", 1) + b"\n")
            w(b"$EndMeshFormat\n$Nodes\n%d\n" % n)
            if binary:
                rec = np.zeros(n, dtype=[("t", "<i4"), ("x", "<f8", 3)])
                rec["t"], rec["x"] = tags[order], nodes[order]
                w(rec.tobytes() + b"\n")
            else:
                w("".join(f"{tags[i]} {x!r} {y!r} {z!r}\n"
                          for i, (x, y, z) in zip(order, nodes[order].tolist())).encode())
            w(b"$EndNodes\n$Elements\n")
            ne = sum(len(b[1]) for b in blocks)
            w(b"%d\n" % ne)
            et = 1
            for etype, c in blocks:
                if binary:
                    w(struct.pack("<3i", etype, len(c), 2))
                    rows = np.column_stack([np.arange(et, et + len(c)), np.full(len(c), 7), np.full(len(c), 1), c])
                    w(rows.astype("<i4").tobytes())
                else:
                    for row in c:
                        w((f"{et} {etype} 2 7 1 " + " ".join(map(str, row)) + "\n").encode())
                        et += 1
                    continue
                et += len(c)
            if binary:
                w(b"\n")
            w(b"$EndElements\n")
    return tags, order


@pytest.mark.parametrize("version,binary", [(4, False), (4, True), (2, False), (2, True)])
def test_gmsh_reader_maps_type11_to_c3d10(tmp_path, version, binary):
    from wnfea_amd.mesh.gmsh_reader import read_msh
    nodes, elems = _box()
    path = tmp_path / f"box_v{version}_{'bin' if binary else 'ascii'}.msh"
    tags, order = _write(path, nodes, _to_gmsh(nodes, elems), version, binary)
    coords, conn = read_msh(path)
    # nodes come back in file order
    np.testing.assert_array_equal(coords, nodes[order])
    assert conn.shape == elems.shape
    # C3D10 mid-edge slots sit at the midpoints of the Abaqus edges
    for s, (a, b) in enumerate(C3D10_EDGES):
        np.testing.assert_allclose(coords[conn[:, 4 + s]], 0.5 * (coords[conn[:, a]] + coords[conn[:, b]]),
                                   rtol=0, atol=1e-15)
    # the same elements as generated, and only the type-11 ones
    inv = np.empty(len(order), dtype=np.int64)
    inv[order] = np.arange(len(order))
    np.testing.assert_array_equal(conn, inv[elems])


def test_gmsh_reader_jacobians_positive(tmp_path):
    from oracle import oracle as O
    from wnfea_amd.mesh.gmsh_reader import read_msh
    nodes, elems = _box()
    path = tmp_path / "box.msh"
    _write(path, nodes, _to_gmsh(nodes, elems), 4, False)
    coords, conn = read_msh(path)
    for el in conn[:12]:
        Ke = O.element_stiffness(coords[el], 210e9, 0.3)
        w = np.linalg.eigvalsh(0.5 * (Ke + Ke.T))
        assert np.sum(w > 1e-8 * w.max()) == 24  # 6 rigid modes, 24 positive


def test_gmsh_reader_errors(tmp_path):
    from wnfea_amd.mesh.gmsh_reader import GmshReadError, read_msh
    p = tmp_path / "bad.msh"
    p.write_bytes(b"$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1 1 1 1\n0 1 0 1\n1\n0 0 0\n$EndNodes\n"
                  b"$Elements\n1 1 1 1\n0 1 15 1\n1 1\n$EndElements\n")
    with pytest.raises(GmshReadError, match="type 11"):
        read_msh(p)
    q = tmp_path / "missing.msh"
    q.write_bytes(b"$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n1\n1 0 0 0\n$EndNodes\n$Elements\n1\n"
                  b"1 11 2 0 1 1 2 3 4 5 6 7 8 9 10\n$EndElements\n")
    with pytest.raises(GmshReadError, match="absent"):
        read_msh(q)


def test_load_msh_populates_model(tmp_path):
    from wnfea_amd.mesh.gmsh_reader import get_nodes_on_plane, load_msh
    from wnfea_amd.model import FEAModel
    nodes, elems = _box()
    path = tmp_path / "box.msh"
    _write(path, nodes, _to_gmsh(nodes, elems), 4, False)
    model = FEAModel()
    model.solid_materials[999] = "stale"
    n, e, m = load_msh(path, model, material_name="S355")
    assert m is model and m.source_file == str(path)
    assert set(m.solid_materials) == set(range(len(e))) and set(m.solid_materials.values()) == {"S355"}
    mat = m.materials["S355"]
    assert (mat.youngs_modulus, mat.poissons_ratio, mat.yield_strength) == (210.0e9, 0.3, 250.0e6)
    assert len(get_nodes_on_plane(n, "x", 0.0)) == len(np.flatnonzero(np.abs(nodes[:, 0]) < 1e-9))


def _parse_vtu(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    arrays = {}
    for da in piece.iter("DataArray"):
        kind = da.get("type")
        vals = np.array(da.text.split(), dtype=np.float64 if kind.startswith("Float") else np.int64)
        nc = int(da.get("NumberOfComponents", "1"))
        arrays[da.get("Name")] = vals.reshape(-1, nc) if nc > 1 else vals
    return root, piece, arrays


def test_vtu_export_geometry_and_fields(tmp_path):
    from wnfea_amd.model import FEAModel
    from wnfea_amd.results import export_vtu
    nodes, elems = _box()
    m = FEAModel()
    m.mesh_nodes = nodes
    m.solid_elements = elems
    m.mesh_elements = np.array([[0, 1], [1, 2]])          # two beams as well
    rng = np.random.default_rng(5)
    U = rng.standard_normal(6 * len(nodes)) * 1e-3
    m.displacements = U
    path = export_vtu(m, tmp_path / "out" / "box.vtu", compute_stresses=False)
    root, piece, a = _parse_vtu(path)
    assert root.get("type") == "UnstructuredGrid" and root.get("header_type") == "UInt64"
    assert int(piece.get("NumberOfPoints")) == len(nodes)
    assert int(piece.get("NumberOfCells")) == len(elems) + 2
    np.testing.assert_array_equal(a["types"], [3, 3] + [24] * len(elems))
    np.testing.assert_array_equal(a["connectivity"], np.concatenate([[0, 1, 1, 2], elems.ravel()]))
    np.testing.assert_array_equal(a["offsets"], np.concatenate([[2, 4], 4 + 10 * np.arange(1, len(elems) + 1)]))
    np.testing.assert_array_equal(a["CellID"], np.arange(len(elems) + 2))
    np.testing.assert_allclose(a["Points"], nodes, rtol=1e-8, atol=1e-300)
    U6 = U.reshape(-1, 6)
    np.testing.assert_allclose(a["Displacement"], U6[:, :3], rtol=1e-8)
    np.testing.assert_allclose(a["Rotation"], U6[:, 3:], rtol=1e-8)
    np.testing.assert_allclose(a["DisplacementMagnitude"], np.linalg.norm(U6[:, :3], axis=1), rtol=1e-8)
    assert "VonMises" not in a and "VonMises_Nodal" not in a
    # every float as %.8e on one line per array
    txt = open(path).read()
    assert f"{U6[0, 0]:.8e}" in txt


def test_vtu_export_three_dof_field_has_no_rotation(tmp_path):
    from wnfea_amd.model import FEAModel
    nodes, elems = _box()
    m = FEAModel()
    m.mesh_nodes = nodes
    m.solid_elements = elems
    m.displacements = np.arange(3 * len(nodes), dtype=np.float64)
    path = m.export_vtu(tmp_path / "b.vtu", compute_stresses=False)
    _, _, a = _parse_vtu(path)
    assert "Rotation" not in a
    np.testing.assert_allclose(a["Displacement"].ravel(), m.displacements, rtol=1e-8)
    m2 = FEAModel()
    m2.mesh_nodes = nodes
    with pytest.raises(ValueError, match="no beam or solid"):
        m2.export_vtu(tmp_path / "c.vtu")
