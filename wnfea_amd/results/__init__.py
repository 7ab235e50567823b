"""Result export (``wnfea/results``)."""
from .vtu_exporter import VTK_LINE, VTK_QUADRATIC_TETRA, VTUExporter, export_vtu

__all__ = ["VTK_LINE", "VTK_QUADRATIC_TETRA", "VTUExporter", "export_vtu"]
