"""ParaView VTU export of solved models (drop-in for
``wnfea/results/vtu_exporter.py:30-269``).

Same file as the reference writes: an XML ``UnstructuredGrid`` with ASCII
``DataArray``s, beams as ``VTK_LINE`` (3) cells first, then C3D10 solids as
``VTK_QUADRATIC_TETRA`` (24) cells (the Abaqus C3D10 node order is VTK's),
point data ``Displacement``, ``DisplacementMagnitude``, ``Rotation`` (only
when a 6-DOF field has a nonzero rotation), ``VonMises_Nodal``, cell data
``VonMises``, ``StressTensor`` [sxx, syy, szz, sxy, syz, szx] and ``CellID``,
every float as ``%.8e`` and every array on one line.  The solid stresses come
from the device stress kernel (Gauss-point stresses averaged per element,
``solid_solver.compute_c3d10_stress_field``); nodal von Mises is the mean of
the von Mises of the cells that use the node, beams included
(``vtu_exporter.py:130-175``).  Text is produced with vectorised formatting,
so a 1 M-DOF model writes in seconds.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from ..model import FEAModel

VTK_LINE = 3
VTK_QUADRATIC_TETRA = 24


def _floats(a: np.ndarray) -> str:
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if a.size == 0:
        return ""
    return " ".join(np.char.mod("%.8e", a).tolist())


def _ints(a: np.ndarray) -> str:
    a = np.asarray(a, dtype=np.int64).ravel()
    return " ".join(map(str, a.tolist()))


def _array(lines: list, kind: str, name: str, text: str, ncomp: int | None = 1, indent: str = "        "):
    comp = f' NumberOfComponents="{ncomp}"' if ncomp is not None else ""
    lines.append(f'{indent}<DataArray type="{kind}" Name="{name}"{comp} format="ascii">')
    lines.append(indent + "  " + text)
    lines.append(f"{indent}</DataArray>")


class VTUExporter:
    """Writes FEAModel instances as .vtu files (``vtu_exporter.py:30-40``)."""

    def __init__(self, precision: int = 8):
        self.precision = precision  # kept for API parity; floats are written as %.8e like the reference

    def export(self, model: FEAModel, filepath, compute_stresses: bool = True) -> str:
        if model.mesh_nodes is None or len(model.mesh_nodes) == 0:
            raise ValueError("FEAModel contains no mesh nodes. Mesh model before exporting.")
        nodes = np.asarray(model.mesh_nodes, dtype=np.float64)
        n_points = len(nodes)
        beams = (np.asarray(model.mesh_elements, dtype=np.int64).reshape(-1, 2)
                 if model.mesh_elements is not None and len(model.mesh_elements) else np.zeros((0, 2), np.int64))
        solids = (np.asarray(model.solid_elements, dtype=np.int64).reshape(-1, 10)
                  if getattr(model, "solid_elements", None) is not None and len(model.solid_elements)
                  else np.zeros((0, 10), np.int64))
        nb, ns = len(beams), len(solids)
        n_cells = nb + ns
        if n_cells == 0:
            raise ValueError("FEAModel has mesh nodes but no beam or solid elements to export.")
        connectivity = np.concatenate([beams.ravel(), solids.ravel()])
        offsets = np.concatenate([2 * np.arange(1, nb + 1), 2 * nb + 10 * np.arange(1, ns + 1)])
        types = np.concatenate([np.full(nb, VTK_LINE), np.full(ns, VTK_QUADRATIC_TETRA)])

        # point data (vtu_exporter.py:100-120): 6 DOF/node, 3 DOF/node, or a
        # zero-padded prefix of a 6 DOF/node layout for other lengths
        u_trans = np.zeros((n_points, 3))
        u_rot = np.zeros((n_points, 3))
        has_u = model.displacements is not None
        if has_u:
            U = np.asarray(model.displacements, dtype=np.float64).reshape(-1)
            if len(U) == 6 * n_points:
                u_trans, u_rot = U.reshape(n_points, 6)[:, :3], U.reshape(n_points, 6)[:, 3:]
            elif len(U) == 3 * n_points:
                u_trans = U.reshape(n_points, 3)
            else:
                tmp = np.zeros(6 * n_points)
                k = min(len(U), 6 * n_points)
                tmp[:k] = U[:k]
                u_trans, u_rot = tmp.reshape(n_points, 6)[:, :3], tmp.reshape(n_points, 6)[:, 3:]
        u_mag = np.linalg.norm(u_trans, axis=1)

        # cell data (vtu_exporter.py:124-175)
        cell_vm = np.zeros(n_cells)
        cell_sig = np.zeros((n_cells, 6))
        nodal_vm = np.zeros(n_points)
        if has_u and compute_stresses:
            if ns:
                from ..solver.solid_solver import compute_c3d10_stress_field
                cs, cv, _ = compute_c3d10_stress_field(model, np.ascontiguousarray(u_trans).reshape(-1))
                cell_sig[nb:] = cs
                cell_vm[nb:] = cv
            for e in range(nb):
                er = (model.element_results or {}).get(e)
                if er is not None:
                    cell_vm[e] = getattr(er, "max_von_mises", 0.0)
                    cell_sig[e, 0] = getattr(er, "axial_stress", 0.0)
            # nodal mean over incident cells, beams and solids alike
            acc = np.bincount(connectivity, weights=np.repeat(cell_vm, np.diff(np.concatenate([[0], offsets]))),
                              minlength=n_points)
            cnt = np.bincount(connectivity, minlength=n_points)
            nz = cnt > 0
            nodal_vm[nz] = acc[nz] / cnt[nz]

        lines = ['<?xml version="1.0"?>',
                 '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">',
                 "  <UnstructuredGrid>",
                 f'    <Piece NumberOfPoints="{n_points}" NumberOfCells="{n_cells}">',
                 "      <PointData>"]
        _array(lines, "Float64", "Displacement", _floats(u_trans), 3)
        _array(lines, "Float64", "DisplacementMagnitude", _floats(u_mag), 1)
        if has_u and np.any(u_rot != 0.0):
            _array(lines, "Float64", "Rotation", _floats(u_rot), 3)
        if has_u and compute_stresses:
            _array(lines, "Float64", "VonMises_Nodal", _floats(nodal_vm), 1)
        lines += ["      </PointData>", "      <CellData>"]
        if has_u and compute_stresses:
            _array(lines, "Float64", "VonMises", _floats(cell_vm), 1)
            _array(lines, "Float64", "StressTensor", _floats(cell_sig), 6)
        _array(lines, "Int64", "CellID", _ints(np.arange(n_cells)), 1)
        lines += ["      </CellData>", "      <Points>"]
        _array(lines, "Float64", "Points", _floats(nodes), 3)
        lines += ["      </Points>", "      <Cells>"]
        _array(lines, "Int64", "connectivity", _ints(connectivity), None)
        _array(lines, "Int64", "offsets", _ints(offsets), None)
        _array(lines, "UInt8", "types", _ints(types), None)
        lines += ["      </Cells>", "    </Piece>", "  </UnstructuredGrid>", "</VTKFile>"]
        out = Path(filepath).resolve()
        out.parent.mkdir(parents=True, exist_ok=True)
        with open(out, "w", encoding="utf-8") as f:
            f.write("\n".join(lines) + "\n")
        return str(out)


def export_vtu(model: FEAModel, filepath, compute_stresses: bool = True) -> str:
    """``vtu_exporter.py:262-269``."""
    return VTUExporter().export(model, filepath, compute_stresses=compute_stresses)
