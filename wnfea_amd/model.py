"""Model carrier of the solve-path API.

``FEAModel`` keeps the reference's field names and meanings
(``wnfea/model.py:45-83``): ``mesh_nodes (N,3) f64``, ``solid_elements (M,10)``
in Abaqus C3D10 order, ``solid_materials {elem_id: material_name}``,
``materials {name: MaterialDef}``, ``supports``/``loads`` lists, beam
``mesh_elements (E,2)`` with ``element_properties``, and the 6-DOF
``displacements (N*6,)`` result slot.  Geometry-import and GUI stages of the
reference are not part of this package.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .boundary.conditions import LoadDef, SupportDef
from .properties.materials import MaterialDef


@dataclass
class PropertyAssignment:
    """Material and section names of one beam element (``model.py:37-41``)."""
    material_name: str
    section_name: str


@dataclass
class FEAModel:
    source_file: str = ""
    materials: dict[str, MaterialDef] = field(default_factory=dict)
    sections: dict = field(default_factory=dict)
    mesh_nodes: np.ndarray | None = None
    mesh_elements: np.ndarray | None = None
    element_properties: dict[int, PropertyAssignment] = field(default_factory=dict)
    solid_elements: np.ndarray | None = None
    solid_materials: dict[int, str] = field(default_factory=dict)
    supports: list[SupportDef] = field(default_factory=list)
    loads: list[LoadDef] = field(default_factory=list)
    couplings: list = field(default_factory=list)               # RigidCoupling (mixed beam-solid models)
    geometry_to_mesh_node_map: dict = field(default_factory=dict)
    displacements: np.ndarray | None = None
    element_results: dict | None = None
    reaction_forces: np.ndarray | None = None

    def validate_for_solving(self) -> list[str]:
        """Same checks as ``model.py:130-145``."""
        errs = []
        if self.mesh_nodes is None or (self.mesh_elements is None and self.solid_elements is None):
            errs.append("No mesh generated. Run the mesher first.")
        if not self.supports:
            errs.append("No supports defined. Add at least one boundary condition.")
        if not self.loads:
            errs.append("No loads defined. Add at least one load.")
        if self.supports and not any(s.has_any_constraint for s in self.supports):
            errs.append("No support has any constrained DOFs.")
        if self.loads and not any(ld.has_any_load for ld in self.loads):
            errs.append("All loads are zero. Add at least one nonzero load.")
        return errs

    def clear_results(self) -> None:
        self.displacements = None
        self.element_results = None
        self.reaction_forces = None

    def clear_mesh(self) -> None:
        """Drop the mesh and everything downstream; geometry-node supports
        and loads stay (``model.py:176-188``)."""
        self.mesh_nodes = None
        self.mesh_elements = None
        self.solid_elements = None
        self.solid_materials.clear()
        self.element_properties.clear()
        self.couplings.clear()
        self.geometry_to_mesh_node_map.clear()
        self.supports = [s for s in self.supports if s.is_geometry_node]
        self.loads = [ld for ld in self.loads if ld.is_geometry_node]
        self.clear_results()

    def export_vtu(self, filepath, compute_stresses: bool = True, precision: int = 8) -> str:
        """ParaView .vtu of this model (``model.py:191-211``)."""
        from .results.vtu_exporter import VTUExporter
        return VTUExporter(precision=precision).export(self, filepath, compute_stresses=compute_stresses)

    # ---- vectorised views used by the device path (not in the reference) ----
    def element_moduli(self, default_E: float = 2.1e11, default_nu: float = 0.3) -> tuple[np.ndarray, np.ndarray]:
        """Per-element (E, nu) arrays with the reference's lookup rule
        (``solid_solver.py:67-80``): the element's named material if present,
        else the first material in ``materials``, else steel defaults.
        A model may carry a precomputed ``solid_E``/``solid_nu`` pair of arrays
        (per-element moduli for sensitivity studies); those win when present."""
        m = 0 if self.solid_elements is None else len(self.solid_elements)
        pre_E = getattr(self, "solid_E", None)
        pre_nu = getattr(self, "solid_nu", None)
        if pre_E is not None:
            E = np.ascontiguousarray(pre_E, dtype=np.float64)
            nu = (np.ascontiguousarray(pre_nu, dtype=np.float64) if pre_nu is not None
                  else np.full(m, self._default_material(default_E, default_nu)[1]))
            return E, nu
        E0, nu0 = self._default_material(default_E, default_nu)
        E = np.full(m, E0, dtype=np.float64)
        nu = np.full(m, nu0, dtype=np.float64)
        if self.solid_materials:
            eids = np.fromiter(self.solid_materials.keys(), dtype=np.int64, count=len(self.solid_materials))
            names = list(self.solid_materials.values())
            for name in set(names):
                mat = self.materials.get(name)
                if mat is None:
                    continue
                sel = eids[np.fromiter((n == name for n in names), dtype=bool, count=len(names))]
                sel = sel[(sel >= 0) & (sel < m)]
                E[sel] = mat.youngs_modulus
                nu[sel] = mat.poissons_ratio
        return E, nu

    def _default_material(self, default_E: float, default_nu: float) -> tuple[float, float]:
        if self.materials:
            mat = next(iter(self.materials.values()))
            return float(mat.youngs_modulus), float(mat.poissons_ratio)
        return default_E, default_nu

    def fixed_dof_mask(self, dofs_per_node: int = 3) -> np.ndarray:
        """Boolean mask of FIXED DOFs over ``n_nodes*dofs_per_node`` (mesh
        nodes only, first ``dofs_per_node`` constraints; ``solid_solver.py:96-103``)."""
        from .boundary.conditions import DOFType
        n = len(self.mesh_nodes)
        mask = np.zeros(n * dofs_per_node, dtype=bool)
        for s in self.supports:
            nid = s.node_id
            if nid < n:
                for d, c in enumerate(s.constraints[:dofs_per_node]):
                    if c.dof_type == DOFType.FIXED:
                        mask[nid * dofs_per_node + d] = True
        return mask

    def fixed_dof_list(self, dofs_per_node: int = 3) -> list[int]:
        """Fixed DOFs in the reference's order: support by support, DOF by DOF."""
        from .boundary.conditions import DOFType
        n = len(self.mesh_nodes)
        out: list[int] = []
        for s in self.supports:
            nid = s.node_id
            if nid < n:
                for d, c in enumerate(s.constraints[:dofs_per_node]):
                    if c.dof_type == DOFType.FIXED:
                        out.append(nid * dofs_per_node + d)
        return out

    def load_vector(self, dofs_per_node: int = 3) -> np.ndarray:
        """Nodal force vector; loads on the same node add (``solid_solver.py:86-93``)."""
        n = len(self.mesh_nodes)
        F = np.zeros(n * dofs_per_node, dtype=np.float64)
        for ld in self.loads:
            nid = ld.node_id
            if nid < n:
                comps = ld.force_vector[:dofs_per_node]
                for d, v in enumerate(comps):
                    F[nid * dofs_per_node + d] += v
        return F
