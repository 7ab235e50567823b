"""Boundary-condition carriers of the solve-path API.

Same names, fields and semantics as the reference's
``wnfea/boundary/conditions.py:19-162`` (per-DOF FREE/FIXED/PRESCRIBED
constraints on six DOFs [Ux, Uy, Uz, Rx, Ry, Rz]; point loads with six
components), so models built for the reference can be handed to this package
unchanged.  Solid (C3D10) solves read only the first three entries.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from enum import Enum

_DOF_LABELS = ("Ux", "Uy", "Uz", "Rx", "Ry", "Rz")


class DOFType(Enum):
    """Constraint kind of one degree of freedom (``conditions.py:19-23``)."""
    FREE = "free"
    FIXED = "fixed"
    PRESCRIBED = "prescribed"


@dataclass
class DOFConstraint:
    """One DOF's constraint; ``value`` only matters when PRESCRIBED."""
    dof_type: DOFType = DOFType.FREE
    value: float = 0.0


@dataclass
class SupportDef:
    """Support on one node (``conditions.py:43-102``)."""
    node_id: int
    is_geometry_node: bool = True
    ux: DOFConstraint = field(default_factory=DOFConstraint)
    uy: DOFConstraint = field(default_factory=DOFConstraint)
    uz: DOFConstraint = field(default_factory=DOFConstraint)
    rx: DOFConstraint = field(default_factory=DOFConstraint)
    ry: DOFConstraint = field(default_factory=DOFConstraint)
    rz: DOFConstraint = field(default_factory=DOFConstraint)
    label: str = ""

    @property
    def constraints(self) -> list[DOFConstraint]:
        return [self.ux, self.uy, self.uz, self.rx, self.ry, self.rz]

    @property
    def has_any_constraint(self) -> bool:
        return any(c.dof_type != DOFType.FREE for c in self.constraints)

    @property
    def fixed_dof_indices(self) -> list[int]:
        return [i for i, c in enumerate(self.constraints) if c.dof_type == DOFType.FIXED]

    @property
    def prescribed_dofs(self) -> dict[int, float]:
        return {i: c.value for i, c in enumerate(self.constraints)
                if c.dof_type == DOFType.PRESCRIBED}

    def __repr__(self) -> str:
        tags = []
        for name, c in zip(_DOF_LABELS, self.constraints):
            if c.dof_type == DOFType.FIXED:
                tags.append(f"{name}=FIXED")
            elif c.dof_type == DOFType.PRESCRIBED:
                tags.append(f"{name}={c.value}")
        where = "geom" if self.is_geometry_node else "mesh"
        return f"SupportDef({where}_node={self.node_id}, {', '.join(tags) or 'all FREE'})"


@dataclass
class LoadDef:
    """Point load on one node (``conditions.py:105-146``)."""
    node_id: int
    is_geometry_node: bool = True
    fx: float = 0.0
    fy: float = 0.0
    fz: float = 0.0
    mx: float = 0.0
    my: float = 0.0
    mz: float = 0.0
    label: str = ""

    @property
    def force_vector(self) -> list[float]:
        return [self.fx, self.fy, self.fz, self.mx, self.my, self.mz]

    @property
    def has_any_load(self) -> bool:
        return any(abs(v) > 1e-12 for v in self.force_vector)


def create_fixed_support(node_id: int, is_geometry_node: bool = True, label: str = "") -> SupportDef:
    """Fully clamped support, all six DOFs FIXED (``conditions.py:149-162``)."""
    def fixed() -> DOFConstraint:
        return DOFConstraint(DOFType.FIXED)
    return SupportDef(node_id=node_id, is_geometry_node=is_geometry_node,
                      ux=fixed(), uy=fixed(), uz=fixed(), rx=fixed(), ry=fixed(), rz=fixed(),
                      label=label)
