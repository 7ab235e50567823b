"""Nonlinear solve entry points (drop-in for ``wnfea/solver/nonlinear_solver.py``).

``solve_nonlinear_jfnk`` is the device Newton solver (``beam_solver``).
``pcg_solve`` is the reference's host PCG driver for a LinearOperator and a
preconditioner (``nonlinear_solver.py:51-156``), with the same tri-precision
bookkeeping: it calls ``apply_adaptive`` on preconditioners that have it
(the FP16 V-cycle while the relative residual is above their switch
tolerance).  Its operator and preconditioner products run on the device
(``MatrixFreeJFNKOperator``, ``BlockBeamAMGPreconditioner``); the Krylov
vectors are the caller's numpy arrays in ``working_dtype``.
"""
from __future__ import annotations

import numpy as np

from .beam_solver import NonLinearConvergenceError, solve_nonlinear_jfnk  # noqa: F401


def _precondition(M, r, rel, counts):
    if M is None:
        return r.copy()
    if hasattr(M, "apply_adaptive"):
        z = M.apply_adaptive(r.astype(np.float64), current_res_rel=rel)
        counts["fp16_iters" if getattr(M, "last_precision_used", "") == "float16" else "fp32_iters"] += 1
        return z
    counts["fp32_iters"] += 1
    return M.matvec(r.astype(np.float64))


def pcg_solve(A_op, b, M_prec=None, tol: float = 1e-6, max_iter: int = 150, working_dtype=np.float64,
              stats: dict | None = None):
    """PCG from x = 0 until ||r|| < tol ||b||.  Returns (x (float64),
    iterations, final ||r||)."""
    wd = np.dtype(working_dtype)
    counts = {"fp16_iters": 0, "fp32_iters": 0}
    r = np.asarray(b, dtype=wd).copy()
    x = np.zeros_like(r)
    nb = float(np.linalg.norm(r))
    if nb < 1e-15:
        if stats is not None:
            stats.update(counts)
        return x.astype(np.float64), 0, 0.0
    z = np.asarray(_precondition(M_prec, r, 1.0, counts), dtype=wd)
    p = z.copy()
    rz = float(r @ z)
    for it in range(max_iter):
        q = np.asarray(A_op.matvec(p.astype(np.float64)), dtype=wd)
        pq = float(p @ q)
        if abs(pq) < 1e-20:
            break
        a = rz / pq
        x += a * p
        r -= a * q
        rn = float(np.linalg.norm(r))
        if rn / (nb + 1e-30) < tol:
            if stats is not None:
                stats.update(counts)
            return x.astype(np.float64), it + 1, rn
        z = np.asarray(_precondition(M_prec, r, rn / (nb + 1e-30), counts), dtype=wd)
        rz_new = float(r @ z)
        p = z + (rz_new / (rz + 1e-30)) * p
        rz = rz_new
    if stats is not None:
        stats.update(counts)
    return x.astype(np.float64), max_iter, float(np.linalg.norm(r))
