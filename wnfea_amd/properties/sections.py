"""Beam cross-section carrier: the subset of ``wnfea/properties/sections.py``
the co-rotational lattice needs (``SectionDef`` fields and
``create_solid_circle``, ``sections.py:109-140``)."""
from __future__ import annotations

import math
from dataclasses import dataclass


@dataclass
class SectionDef:
    name: str
    area: float
    iy: float
    iz: float
    j: float
    outer_radius: float = 0.0
    inner_radius: float = 0.0


def create_solid_circle(name: str, radius: float) -> SectionDef:
    if radius <= 0:
        raise ValueError(f"Radius must be positive, got {radius}")
    a = math.pi * radius ** 2
    i = math.pi / 4.0 * radius ** 4
    return SectionDef(name=name, area=a, iy=i, iz=i, j=2.0 * i,
                      outer_radius=radius, inner_radius=0.0)
