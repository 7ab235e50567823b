"""Isotropic material carrier (``wnfea/properties/materials.py:11-47``)."""
from __future__ import annotations

from dataclasses import dataclass


@dataclass
class MaterialDef:
    name: str
    youngs_modulus: float
    poissons_ratio: float
    yield_strength: float

    @property
    def shear_modulus(self) -> float:
        return self.youngs_modulus / (2.0 * (1.0 + self.poissons_ratio))

    def validate(self) -> list[str]:
        errs = []
        if self.youngs_modulus <= 0:
            errs.append(f"Young's Modulus must be positive, got {self.youngs_modulus}")
        if not 0.0 < self.poissons_ratio < 0.5:
            errs.append(f"Poisson's Ratio must be in (0, 0.5), got {self.poissons_ratio}")
        if self.yield_strength <= 0:
            errs.append(f"Yield Strength must be positive, got {self.yield_strength}")
        return errs
