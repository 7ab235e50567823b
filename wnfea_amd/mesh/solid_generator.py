"""Structured C3D10 mesh builders.

``generate_c3d10_structured_block`` keeps the reference signature and output
conventions of ``wnfea/mesh/solid_generator.py:21-213``:

* corner nodes first, in (i, j, k) lexicographic grid order, then one midside
  node per distinct tet edge in first-seen order over (element, edge), with
  the edge order (0-1, 1-2, 2-0, 0-3, 1-3, 2-3) (``:125-157``);
* hex corners c0..c7 = (i,j,k), (i+1,j,k), (i+1,j+1,k), (i,j+1,k), then the
  same four at k+1 (``:92-101``); a tet with negative signed volume gets its
  vertices 1 and 2 swapped (``:109-120``);
* root face x = 0 fully clamped, total tip load split equally over the nodes
  of the face x = length (``:177-196``).

Two options extend the reference: ``tets_per_hex=6`` selects the 6-tet Kuhn
split (all six tets share the c0-c6 diagonal, which makes the split conforming
on any structured grid) and ``load_direction`` picks the loaded axis (the
reference loads +y).  ``generate_c3d10_ibeam`` builds the structured I-beam
extrusion on the same machinery with a cell mask over the section's tensor
grid.  Everything is vectorised numpy; no per-element Python loop.
"""
from __future__ import annotations

import time

import numpy as np

from ..boundary.conditions import LoadDef, create_fixed_support
from ..model import FEAModel
from ..properties.materials import MaterialDef

# 5-tet split with parity alternation (solid_generator.py:68-81)
_HEX_TETS_EVEN = np.array([(0, 1, 3, 4), (1, 2, 3, 6), (1, 4, 5, 6), (3, 4, 6, 7), (1, 3, 4, 6)], dtype=np.int64)
_HEX_TETS_ODD = np.array([(0, 1, 2, 5), (0, 2, 3, 7), (0, 4, 5, 7), (2, 5, 6, 7), (0, 2, 5, 7)], dtype=np.int64)
# 6-tet Kuhn split: one tet per axis ordering of the monotone path c0 -> c6
_HEX_TETS_KUHN = np.array([(0, 1, 2, 6), (0, 1, 5, 6), (0, 3, 2, 6), (0, 3, 7, 6), (0, 4, 5, 6), (0, 4, 7, 6)],
                          dtype=np.int64)
_EDGES = np.array([(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)], dtype=np.int64)
_HEX_OFFSETS = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)],
                        dtype=np.int64)


def structured_c3d10_mesh(x_grid, y_grid, z_grid, cell_mask=None, tets_per_hex: int = 5):
    """C3D10 connectivity and coordinates on a (masked) tensor grid.

    Returns ``(coords (N,3) f64, elements (M,10) int32)``.  Unused grid corners
    (outside the mask) are dropped; the kept corners keep their lexicographic
    order, so an unmasked grid numbers exactly like the reference generator.
    """
    x_grid = np.asarray(x_grid, dtype=np.float64)
    y_grid = np.asarray(y_grid, dtype=np.float64)
    z_grid = np.asarray(z_grid, dtype=np.float64)
    nx, ny, nz = len(x_grid) - 1, len(y_grid) - 1, len(z_grid) - 1
    if min(nx, ny, nz) < 1:
        raise ValueError("each grid direction needs at least one cell")
    if tets_per_hex not in (5, 6):
        raise ValueError("tets_per_hex must be 5 or 6")
    if cell_mask is None:
        cells = np.indices((nx, ny, nz)).reshape(3, -1).T
    else:
        cell_mask = np.asarray(cell_mask, dtype=bool)
        if cell_mask.shape != (nx, ny, nz):
            raise ValueError(f"cell_mask shape {cell_mask.shape} != {(nx, ny, nz)}")
        cells = np.argwhere(cell_mask)
    X, Y, Z = np.meshgrid(x_grid, y_grid, z_grid, indexing="ij")
    corner_xyz = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()])
    sj, si = nz + 1, (ny + 1) * (nz + 1)

    corner_ijk = cells[:, None, :] + _HEX_OFFSETS[None, :, :]              # (H, 8, 3)
    hex_corner = corner_ijk[..., 0] * si + corner_ijk[..., 1] * sj + corner_ijk[..., 2]
    if tets_per_hex == 6:
        tets = hex_corner[:, _HEX_TETS_KUHN]                                # (H, 6, 4)
    else:
        odd = (cells.sum(axis=1) % 2).astype(bool)
        tets = np.where(odd[:, None, None], hex_corner[:, _HEX_TETS_ODD], hex_corner[:, _HEX_TETS_EVEN])
    tets = tets.reshape(-1, 4)

    # orientation: swap v1, v2 where the signed volume is negative (:109-120)
    p = corner_xyz[tets]
    d1, d2, d3 = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]
    det = (d1[:, 0] * (d2[:, 1] * d3[:, 2] - d2[:, 2] * d3[:, 1])
           - d1[:, 1] * (d2[:, 0] * d3[:, 2] - d2[:, 2] * d3[:, 0])
           + d1[:, 2] * (d2[:, 0] * d3[:, 1] - d2[:, 1] * d3[:, 0]))
    neg = det < 0.0
    tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()

    used = np.unique(tets)
    remap = np.full(corner_xyz.shape[0], -1, dtype=np.int64)
    remap[used] = np.arange(len(used))
    tets = remap[tets]
    corners = corner_xyz[used]
    n_c = len(used)

    ea = tets[:, _EDGES[:, 0]].ravel()
    eb = tets[:, _EDGES[:, 1]].ravel()
    lo, hi = np.minimum(ea, eb), np.maximum(ea, eb)
    key = lo * n_c + hi
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")              # first-seen order
    rank = np.empty(len(uniq), dtype=np.int64)
    rank[order] = np.arange(len(uniq))
    mid_id = n_c + rank[inv]
    u_lo, u_hi = uniq[order] // n_c, uniq[order] % n_c
    mids = 0.5 * (corners[u_lo] + corners[u_hi])

    elements = np.empty((tets.shape[0], 10), dtype=np.int32)
    elements[:, :4] = tets
    elements[:, 4:] = mid_id.reshape(-1, 6)
    coords = np.vstack([corners, mids]) if len(mids) else corners
    return np.ascontiguousarray(coords), elements


def _attach_model(coords, elements, E, nu, yield_strength=2.5e8, material="Steel") -> FEAModel:
    model = FEAModel()
    model.mesh_nodes = coords
    model.solid_elements = elements
    model.materials[material] = MaterialDef(name=material, youngs_modulus=E, poissons_ratio=nu,
                                            yield_strength=yield_strength)
    model.solid_materials = dict.fromkeys(range(len(elements)), material)
    return model


def _axis(load_direction: str) -> int:
    try:
        return {"x": 0, "y": 1, "z": 2}[load_direction.lstrip("+-")]
    except KeyError:
        raise ValueError(f"load_direction must be one of x/y/z, got {load_direction!r}") from None


def generate_c3d10_structured_block(
    length: float = 10.0,
    width: float = 1.0,
    height: float = 1.0,
    nx: int = 40,
    ny: int = 4,
    nz: int = 4,
    E: float = 2.1e11,
    nu: float = 0.3,
    tip_load_total: float = 10000.0,
    tets_per_hex: int = 5,
    load_direction: str = "y",
) -> tuple[FEAModel, dict]:
    """Cantilever box of C3D10 tets (reference signature, ``solid_generator.py:21-31``).

    Root face x = 0 clamped; ``tip_load_total`` split equally over the nodes of
    the face x = length along ``load_direction`` (default +y as the reference).
    """
    t0 = time.perf_counter()
    coords, elements = structured_c3d10_mesh(np.linspace(0.0, length, nx + 1),
                                             np.linspace(0.0, width, ny + 1),
                                             np.linspace(0.0, height, nz + 1),
                                             tets_per_hex=tets_per_hex)
    model = _attach_model(coords, elements, E, nu)
    tol = 1e-6 * length
    root = np.where(coords[:, 0] <= tol)[0]
    tip = np.where(coords[:, 0] >= length - tol)[0]
    model.supports = [create_fixed_support(int(n), is_geometry_node=False) for n in root]
    per_node = tip_load_total / max(len(tip), 1)
    ax = _axis(load_direction)
    for n in tip:
        f = [0.0, 0.0, 0.0]
        f[ax] = per_node
        model.loads.append(LoadDef(node_id=int(n), fx=f[0], fy=f[1], fz=f[2], is_geometry_node=False))
    meta = {
        "length": length, "width": width, "height": height, "grid": (nx, ny, nz),
        "tets_per_hex": tets_per_hex,
        "n_nodes": len(coords), "n_elements": len(elements), "n_dofs": 3 * len(coords),
        "root_nodes_fixed": len(root), "tip_nodes_loaded": len(tip),
        "build_time_sec": time.perf_counter() - t0,
    }
    return model, meta


def ibeam_section_grid(flange_width=0.2, depth=0.3, flange_thickness=0.02, web_thickness=0.012,
                       n_flange_out=8, n_web_across=2, n_flange_thick=2, n_web_height=22):
    """Tensor-grid lines and cell mask of an I-section in the (y, z) plane.

    y runs across the flanges (centred on 0), z runs over the depth from 0 to
    ``depth``.  Default counts give 116 quads per section.
    """
    bo = 0.5 * flange_width
    tw = 0.5 * web_thickness
    y = np.concatenate([np.linspace(-bo, -tw, n_flange_out + 1)[:-1],
                        np.linspace(-tw, tw, n_web_across + 1)[:-1],
                        np.linspace(tw, bo, n_flange_out + 1)])
    z = np.concatenate([np.linspace(0.0, flange_thickness, n_flange_thick + 1)[:-1],
                        np.linspace(flange_thickness, depth - flange_thickness, n_web_height + 1)[:-1],
                        np.linspace(depth - flange_thickness, depth, n_flange_thick + 1)])
    ny, nz = len(y) - 1, len(z) - 1
    mask = np.zeros((ny, nz), dtype=bool)
    mask[:, :n_flange_thick] = True
    mask[:, nz - n_flange_thick:] = True
    mask[n_flange_out:n_flange_out + n_web_across, :] = True
    return y, z, mask


def generate_c3d10_ibeam(
    length: float = 3.0,
    flange_width: float = 0.2,
    depth: float = 0.3,
    flange_thickness: float = 0.02,
    web_thickness: float = 0.012,
    n_layers: int = 300,
    n_flange_out: int = 8,
    n_web_across: int = 2,
    n_flange_thick: int = 2,
    n_web_height: int = 22,
    E: float = 2.1e11,
    nu: float = 0.3,
    shear_total: float = -50e3,
    torque: float = 5e3,
    tets_per_hex: int = 6,
) -> tuple[FEAModel, dict]:
    """Structured C3D10 I-beam along x, root (x = 0) clamped.

    Tip face (x = length) loads: ``shear_total`` along z split equally over the
    tip nodes, plus a torsion couple ``torque`` about the x axis through the
    section centroid (0, depth/2), distributed as f_i = T (e_x x r_i) / sum|r_j|^2
    with r_i the in-plane offset from the centroid, so that sum f_i = 0 and
    sum r_i x f_i = T e_x exactly.
    """
    t0 = time.perf_counter()
    y, z, mask2 = ibeam_section_grid(flange_width, depth, flange_thickness, web_thickness,
                                     n_flange_out, n_web_across, n_flange_thick, n_web_height)
    x = np.linspace(0.0, length, n_layers + 1)
    mask3 = np.broadcast_to(mask2[None, :, :], (n_layers, len(y) - 1, len(z) - 1))
    coords, elements = structured_c3d10_mesh(x, y, z, cell_mask=mask3, tets_per_hex=tets_per_hex)
    model = _attach_model(coords, elements, E, nu)
    tol = 1e-6 * length
    root = np.where(coords[:, 0] <= tol)[0]
    tip = np.where(coords[:, 0] >= length - tol)[0]
    model.supports = [create_fixed_support(int(n), is_geometry_node=False) for n in root]
    f = np.zeros((len(tip), 3))
    f[:, 2] = shear_total / max(len(tip), 1)
    if torque != 0.0 and len(tip):
        ry = coords[tip, 1] - 0.0
        rz = coords[tip, 2] - 0.5 * depth
        s = float(np.sum(ry * ry + rz * rz))
        f[:, 1] += torque * (-rz) / s
        f[:, 2] += torque * ry / s
    model.loads = [LoadDef(node_id=int(n), fx=float(a), fy=float(b), fz=float(c), is_geometry_node=False)
                   for n, (a, b, c) in zip(tip, f)]
    meta = {
        "length": length, "n_layers": n_layers, "section_quads": int(mask2.sum()),
        "tets_per_hex": tets_per_hex,
        "n_nodes": len(coords), "n_elements": len(elements), "n_dofs": 3 * len(coords),
        "root_nodes_fixed": len(root), "tip_nodes_loaded": len(tip), "tip_nodes": tip,
        "build_time_sec": time.perf_counter() - t0,
    }
    return model, meta


def ibeam_tip_load_cases(coords: np.ndarray, tip_nodes, n_cases: int = 32, seed: int = 0,
                         force_sigma: float = 20e3, torque_sigma: float = 2e3, centroid=(0.0, 0.15)) -> np.ndarray:
    """Seeded tip load cases on the I-beam, (3 N, n_cases): per case a resultant
    force F ~ N(0, force_sigma) per component split equally over the tip
    nodes, and a moment M ~ N(0, torque_sigma) per component about the section
    centroid: M_x as the torsion couple f = M_x (e_x x r) / sum |r|^2, M_y and
    M_z as axial forces linear over the section, f_x = M_y z' / sum z'^2 and
    f_x = -M_z y' / sum y'^2 (y', z' measured from the centroid).  Each part
    has exactly the stated resultant."""
    rng = np.random.default_rng(seed)
    tip = np.asarray(tip_nodes, dtype=np.int64)
    ry = coords[tip, 1] - centroid[0]
    rz = coords[tip, 2] - centroid[1]
    s_yz, s_y, s_z = float(np.sum(ry * ry + rz * rz)), float(np.sum(ry * ry)), float(np.sum(rz * rz))
    out = np.zeros((3 * len(coords), n_cases))
    for k in range(n_cases):
        F = rng.normal(0.0, force_sigma, 3)
        M = rng.normal(0.0, torque_sigma, 3)
        f = np.tile(F / len(tip), (len(tip), 1))
        f[:, 1] += M[0] * (-rz) / s_yz
        f[:, 2] += M[0] * ry / s_yz
        f[:, 0] += M[1] * rz / s_z - M[2] * ry / s_y
        out[3 * tip, k] = f[:, 0]
        out[3 * tip + 1, k] = f[:, 1]
        out[3 * tip + 2, k] = f[:, 2]
    return out
