"""Cubic space-frame lattices of 2-node beams (6 DOF per node).

Node numbering and element order follow the reference's frame builder
(``test_gpu_amg.py:36-80``): nodes with ix outermost and iz innermost, then
the x-, y- and z-directed members in the same loop order.  The C4 workload is
the 40x40x40-cell frame with 1 m members, circular r = 0.02 m section,
E = 70 GPa, nu = 0.33, base (z = 0) nodes clamped and seeded lateral plus
compressive loads on the top face.
"""
from __future__ import annotations

import numpy as np

from ..boundary.conditions import LoadDef, create_fixed_support
from ..model import FEAModel, PropertyAssignment
from ..properties.materials import MaterialDef
from ..properties.sections import create_solid_circle


def lattice_nodes_and_beams(nx: int, ny: int, nz: int, spacing: float = 1.0):
    """Nodes ((nx+1)(ny+1)(nz+1), 3) and beams (E, 2) of an nx x ny x nz-cell frame."""
    ix, iy, iz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    nodes = np.column_stack([ix.ravel(), iy.ravel(), iz.ravel()]).astype(np.float64) * spacing
    idx = np.arange(len(nodes)).reshape(nx + 1, ny + 1, nz + 1)
    bx = np.column_stack([idx[:-1, :, :].ravel(), idx[1:, :, :].ravel()])
    by = np.column_stack([idx[:, :-1, :].ravel(), idx[:, 1:, :].ravel()])
    bz = np.column_stack([idx[:, :, :-1].ravel(), idx[:, :, 1:].ravel()])
    return nodes, np.vstack([bx, by, bz]).astype(np.int32)


def generate_beam_lattice(n_cells: int = 40, spacing: float = 1.0, radius: float = 0.02, E: float = 70e9,
                          nu: float = 0.33, seed: int = 4, lateral_sigma: float = 1.0e3,
                          compression_mean: float = 5.0e3, compression_sigma: float = 1.0e3):
    """Cubic lattice model: base nodes clamped (6 DOF), every top node loaded
    with fx, fy ~ N(0, lateral_sigma) and fz = -|N(compression_mean,
    compression_sigma)| (seeded)."""
    nodes, beams = lattice_nodes_and_beams(n_cells, n_cells, n_cells, spacing)
    model = FEAModel()
    mat = MaterialDef(name="Aluminium", youngs_modulus=E, poissons_ratio=nu, yield_strength=2.7e8)
    sec = create_solid_circle("lattice_rod", radius)
    model.materials[mat.name] = mat
    model.sections[sec.name] = sec
    model.mesh_nodes = nodes
    model.mesh_elements = beams
    model.element_properties = dict.fromkeys(range(len(beams)), PropertyAssignment(mat.name, sec.name))
    base = np.flatnonzero(nodes[:, 2] <= 1e-9)
    top = np.flatnonzero(nodes[:, 2] >= n_cells * spacing - 1e-9)
    model.supports = [create_fixed_support(int(n), is_geometry_node=False) for n in base]
    rng = np.random.default_rng(seed)
    fxy = rng.normal(0.0, lateral_sigma, size=(len(top), 2))
    fz = -np.abs(rng.normal(compression_mean, compression_sigma, size=len(top)))
    model.loads = [LoadDef(node_id=int(n), fx=float(a), fy=float(b), fz=float(c), is_geometry_node=False)
                   for n, (a, b), c in zip(top, fxy, fz)]
    meta = {"n_cells": n_cells, "n_nodes": len(nodes), "n_beams": len(beams), "n_dofs": 6 * len(nodes),
            "base_nodes": len(base), "top_nodes": len(top), "seed": seed}
    return model, meta


def beam_props(model: FEAModel) -> np.ndarray:
    """(E, G, A, Iy, Iz, J) per beam from the model's property assignments
    (``fast_kernels.py:276-287``)."""
    n = len(model.mesh_elements)
    out = np.zeros((n, 6))
    cache = {}
    for e in range(n):
        a = model.element_properties[e]
        key = (a.material_name, a.section_name)
        if key not in cache:
            m, s = model.materials[a.material_name], model.sections[a.section_name]
            cache[key] = (m.youngs_modulus, m.shear_modulus, s.area, s.iy, s.iz, s.j)
        out[e] = cache[key]
    return out
