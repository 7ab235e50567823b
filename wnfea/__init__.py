"""The reference's import path for the solve path.

Callers written against WNFEA import ``wnfea.solver.solid_solver``,
``wnfea.mesh.rcm``, ``wnfea.model`` and so on (``wnfea/__init__.py:7-31``
and the reference's tests, e.g. ``tests/test_solid_bench.py:12-13``).  This
package makes those imports resolve, unchanged, to the MI355X modules of
``wnfea_amd``: each ``wnfea.<sub>`` name below is the very module object
``wnfea_amd.<sub>`` (one module state, whichever name a caller uses).  Only
the solve path is provided; the reference's CAD, meshing-from-STEP, linear
beam solver, MFEM bridge and post-processing modules are not part of it, and
importing them raises ImportError as for any absent module.
"""
from __future__ import annotations

import importlib
import sys

# reference module (relative to wnfea) -> module of this implementation
_MODULES = {
    "model": "model",
    "boundary": "boundary",
    "boundary.conditions": "boundary.conditions",
    "boundary.coupling": "boundary.coupling",
    "elements": "elements",
    "elements.c3d10": "elements.c3d10",
    "mesh": "mesh",
    "mesh.rcm": "mesh.rcm",
    "mesh.solid_generator": "mesh.solid_generator",
    "properties": "properties",
    "properties.materials": "properties.materials",
    "properties.sections": "properties.sections",
    "results": "results",
    "results.vtu_exporter": "results.vtu_exporter",
    "solver": "solver",
    "solver.amg_preconditioner": "solver.amg_preconditioner",
    "solver.assembler": "solver.assembler",
    "solver.bsr_matrix": "solver.bsr_matrix",
    "solver.corotational_beam": "solver.corotational_beam",
    "solver.dof_manager": "solver.dof_manager",
    "solver.jfnk_operator": "solver.jfnk_operator",
    "solver.mixed_precision": "solver.mixed_precision",
    "solver.nonlinear_solver": "solver.nonlinear_solver",
    "solver.residual": "solver.residual",
    "solver.solid_solver": "solver.solid_solver",
    "solver.fast_kernels": "solver.fast_kernels",
}


def _install() -> None:
    pkg = sys.modules[__name__]
    for ref_name, own in sorted(_MODULES.items(), key=lambda kv: kv[0].count(".")):
        mod = importlib.import_module("wnfea_amd." + own)
        full = f"{__name__}.{ref_name}"
        sys.modules.setdefault(full, mod)
        parent_name, _, leaf = full.rpartition(".")
        parent = pkg if parent_name == __name__ else sys.modules[parent_name]
        setattr(parent, leaf, mod)


_install()

from wnfea_amd.boundary.conditions import (DOFConstraint, DOFType, LoadDef, SupportDef,  # noqa: E402
                                           create_fixed_support)
from wnfea_amd.boundary.coupling import RigidCoupling  # noqa: E402
from wnfea_amd.elements.c3d10 import (b_matrix_c3d10, element_internal_forces_c3d10,  # noqa: E402
                                      element_stiffness_c3d10, element_stresses_c3d10, jacobian_c3d10,
                                      shape_functions_c3d10)
from wnfea_amd.model import FEAModel  # noqa: E402
from wnfea_amd.results.vtu_exporter import VTUExporter  # noqa: E402
from wnfea_amd.solver.amg_preconditioner import BlockBeamAMGPreconditioner  # noqa: E402
from wnfea_amd.solver.dof_manager import DOFManager  # noqa: E402
from wnfea_amd.solver.jfnk_operator import MatrixFreeJFNKOperator  # noqa: E402
from wnfea_amd.solver.mixed_precision import FULL_FP64, MIXED_FP32, TRI_PRECISION, PrecisionConfig  # noqa: E402
from wnfea_amd.solver.nonlinear_solver import NonLinearConvergenceError, solve_nonlinear_jfnk  # noqa: E402

__all__ = [
    "FEAModel", "SupportDef", "LoadDef", "DOFConstraint", "DOFType", "create_fixed_support", "RigidCoupling",
    "element_stiffness_c3d10", "element_internal_forces_c3d10", "element_stresses_c3d10", "b_matrix_c3d10",
    "shape_functions_c3d10", "jacobian_c3d10", "solve_nonlinear_jfnk", "NonLinearConvergenceError",
    "MatrixFreeJFNKOperator", "BlockBeamAMGPreconditioner", "DOFManager", "PrecisionConfig", "MIXED_FP32",
    "TRI_PRECISION", "FULL_FP64", "VTUExporter",
]
