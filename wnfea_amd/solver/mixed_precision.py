"""Precision policy of the solve path (drop-in for
``wnfea/solver/mixed_precision.py:24-122``).

``PrecisionConfig`` keeps the reference's fields, presets and summary.  On
the device the policy maps to the V-cycle value storage of the mixed solver
(the Krylov vectors, the SpMV of K, alpha/beta and the true residual are FP64
in every mode, which is what makes the C3D10 I-beam converge to 1e-8; see
``solid_solver.solve_mixed``):

* ``FULL_FP64``      -> FP64 V-cycle;
* ``MIXED_FP32``     -> FP32 V-cycle;
* ``TRI_PRECISION``  -> FP16-stored V-cycle until the relative residual
  falls below ``switch_tol``, then CG restarts on the FP64 true residual with
  the FP32 V-cycle (the reference switches its preconditioner at the same
  threshold, ``nonlinear_solver.py:60-156``).
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np


@dataclass
class PrecisionConfig:
    inner_dtype: np.dtype = np.float32
    outer_dtype: np.dtype = np.float64
    precond_coarse_dtype: np.dtype = np.float16
    switch_tol: float = 1e-2
    use_tri_precision: bool = False
    enabled: bool = True

    @property
    def inner_eps_scale(self) -> float:
        return float(np.sqrt(np.finfo(self.inner_dtype).eps))

    @property
    def outer_eps_scale(self) -> float:
        return float(np.sqrt(np.finfo(self.outer_dtype).eps))

    @property
    def working_dtype(self) -> np.dtype:
        return np.dtype(self.inner_dtype) if self.enabled else np.dtype(self.outer_dtype)

    @property
    def amg_precision(self) -> str:
        """V-cycle storage of the device solver for the inner dtype."""
        if not self.enabled or np.dtype(self.inner_dtype) == np.float64:
            return "fp64"
        return "fp16" if np.dtype(self.inner_dtype) == np.float16 else "fp32"

    def summary(self) -> str:
        if not self.enabled:
            return f"Uniform {np.dtype(self.outer_dtype).name} (mixed precision disabled)"
        inner, outer = np.dtype(self.inner_dtype).name, np.dtype(self.outer_dtype).name
        if self.use_tri_precision:
            coarse = np.dtype(self.precond_coarse_dtype).name
            return (f"Tri-precision: outer={outer} (residual/accumulation), inner={inner} (PCG/JFNK), "
                    f"preconditioner={coarse}->{inner} (switch at rel_res={self.switch_tol:.1e})")
        return f"Mixed precision: inner={inner} (PCG/JFNK/AMG), outer={outer} (residual/accumulation)"


MIXED_FP32 = PrecisionConfig(inner_dtype=np.float32, outer_dtype=np.float64, enabled=True, use_tri_precision=False)
TRI_PRECISION = PrecisionConfig(inner_dtype=np.float32, outer_dtype=np.float64, precond_coarse_dtype=np.float16,
                                switch_tol=1e-2, use_tri_precision=True, enabled=True)
FULL_FP64 = PrecisionConfig(inner_dtype=np.float64, outer_dtype=np.float64, enabled=False)


def solve_with_policy(system, amg, config: PrecisionConfig, tol: float = 1e-8, max_iter: int = 2000):
    """Solve the resident system to ||F - K u|| <= tol ||F|| under a precision
    policy.  The hierarchy must store the precisions the policy uses.
    Returns (u, iterations, ||F-Ku||/||F||, seconds)."""
    import torch
    from .solid_solver import pcg_device, solve_mixed, true_residual
    t0 = time.perf_counter()
    if not config.use_tri_precision:
        u, it, _, rel, _ = solve_mixed(system, amg, tol=tol, amg_precision=config.amg_precision, max_iter=max_iter)
        return u, it, rel, time.perf_counter() - t0
    # phase 1: FP16-stored cycle down to switch_tol
    u, it1, _, _ = pcg_device(system, rtol=config.switch_tol, max_iter=max_iter, preconditioner="amg",
                              precision="fp64", amg=amg, amg_precision="fp16")
    r = system.F - system.K.spmv(u)
    saved = system.F
    try:
        system.F = r.contiguous()
        fn, rn = float(torch.linalg.norm(saved)), float(torch.linalg.norm(r))
        du, it2, _, _, _ = solve_mixed(system, amg, tol=tol * fn / max(rn, 1e-300), amg_precision="fp32",
                                       max_iter=max_iter)
    finally:
        system.F = saved
    u = u + du
    return u, it1 + it2, true_residual(system, u), time.perf_counter() - t0
