"""Jacobian-free tangent action (drop-in for ``wnfea/solver/jfnk_operator.py:26-134``).

w = J(U) v by a forward difference of the equilibrium residual,
eps = eps_scale (1 + ||U||) / ||v|| with eps_scale = sqrt(machine eps) of
the working precision, the perturbation zeroed on constrained DOFs and the
identity on their rows.  Each product evaluates the residual once: the
internal forces run on the device (``residual.compute_internal_forces``, mesh
kept resident between calls).  The device Newton solver
(``BeamSystem.solve_newton``) uses an assembled tangent instead.
"""
from __future__ import annotations

import numpy as np
from scipy.sparse.linalg import LinearOperator

from .dof_manager import DOFManager
from .residual import compute_equilibrium_residual, get_boundary_constraints


class MatrixFreeJFNKOperator(LinearOperator):
    def __init__(self, model, U, F_ext, load_factor: float = 1.0, R_current=None, dof_mgr: DOFManager | None = None,
                 eps_scale: float | None = None, working_dtype=np.float64):
        self.model = model
        self.dof_mgr = dof_mgr if dof_mgr is not None else DOFManager(model)
        self.working_dtype = np.dtype(working_dtype)
        self.eps_scale = float(np.sqrt(np.finfo(self.working_dtype).eps)) if eps_scale is None else float(eps_scale)
        self.load_factor = float(load_factor)
        self.U = np.asarray(U, dtype=self.working_dtype)
        self.F_ext = np.asarray(F_ext, dtype=self.working_dtype)
        self.constrained_dofs, self.prescribed_vals = get_boundary_constraints(model, self.dof_mgr)
        self.is_constrained = np.zeros(len(self.U), dtype=bool)
        self.is_constrained[self.constrained_dofs] = True
        self.R_current = self._residual(self.U) if R_current is None else np.asarray(R_current, self.working_dtype)
        super().__init__(dtype=np.float64, shape=(len(self.U), len(self.U)))

    def _residual(self, U):
        return compute_equilibrium_residual(self.model, np.asarray(U, np.float64), self.F_ext.astype(np.float64),
                                            self.load_factor, self.constrained_dofs, self.prescribed_vals,
                                            dof_mgr=self.dof_mgr).astype(self.working_dtype)

    def update_state(self, U, R_current=None) -> None:
        """Linearise about a new state (``jfnk_operator.py:85-96``)."""
        self.U = np.asarray(U, dtype=self.working_dtype)
        self.R_current = self._residual(self.U) if R_current is None else np.asarray(R_current, self.working_dtype)

    def _matvec(self, v):
        v = np.asarray(v, dtype=self.working_dtype).ravel()
        vn = float(np.linalg.norm(v))
        if vn < 1e-14:
            return np.zeros(len(v))
        eps = self.eps_scale * (1.0 + float(np.linalg.norm(self.U))) / vn
        vp = np.where(self.is_constrained, 0.0, v).astype(self.working_dtype)
        w = (self._residual(self.U + eps * vp) - self.R_current) / eps
        w[self.is_constrained] = v[self.is_constrained]
        return w.astype(np.float64)
