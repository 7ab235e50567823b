"""Geometrically nonlinear co-rotational beam frames on the device.

Drop-in for the reference's beam path:
* ``compute_internal_forces_fast(model, U_full)`` (``fast_kernels.py:359-448``)
  -> ``wnfea_beam_forces``;
* ``assemble_global_system_sparse(model)`` (linear 12x12 beams, 6-DOF
  system; ``assembler.py:285-450``) -> ``wnfea_beam_ke_linear`` +
  ``wnfea_bsr_assemble_elem``;
* ``solve_nonlinear_jfnk(model, n_load_steps, ...)``
  (``nonlinear_solver.py:159-376``) -> ``BeamSystem.solve_newton``.

The Newton loop keeps the reference's structure -- load steps lambda =
k/n, residual R = F_int(U) - lambda F_ext with R = U on fixed DOFs (FP64),
Eisenstat-Walker inner tolerance eta = min(0.1, 0.5 (|R|/|R_prev|)^2),
backtracking on |R| down to alpha = 1/16 -- but replaces the matrix-free
finite-difference Jacobian action by an assembled tangent: each element's
12x12 tangent is the central difference of its co-rotational forces
(one device thread per element column), symmetrised and assembled into the
6x6-block matrix, and the inner solve is FP64 CG on that tangent with the
FP32 SA-AMG V-cycle (hierarchy built once from the tangent at U = 0, finest
level refreshed with every tangent).
"""
from __future__ import annotations

import time

import numpy as np

from .. import _lib
from ..model import FEAModel


class NonLinearConvergenceError(RuntimeError):
    """Newton did not converge inside a load step (``nonlinear_solver.py:20-22``)."""


class BeamSystem:
    """Device-resident 6-DOF beam frame: geometry, tangent matrix, AMG."""

    def __init__(self, model: FEAModel, sigma: int = 4096, h_rel: float = 0.0):
        import torch
        from ..mesh.lattice import beam_props
        from .bsr_matrix import DeviceBSR
        from .solid_solver import SolidSystem, _sync
        t0 = time.perf_counter()
        L = _lib.load()
        dev = _lib.require_gpu()
        self.model = model
        self.h_rel = h_rel
        nodes = np.ascontiguousarray(model.mesh_nodes, dtype=np.float64)
        conn = np.ascontiguousarray(model.mesh_elements, dtype=np.int32)
        if conn.ndim != 2 or conn.shape[1] != 2:
            raise ValueError("mesh_elements must have shape (E, 2)")
        n, M = len(nodes), len(conn)
        if M and (conn.min() < 0 or conn.max() >= n):
            raise ValueError("mesh_elements reference a node outside mesh_nodes")
        props = np.ascontiguousarray(beam_props(model))
        fixed = model.fixed_dof_mask(6)
        node_fix = (fixed.reshape(n, 6) * (1 << np.arange(6))).sum(axis=1).astype(np.uint8)
        node_free = (node_fix != 63).astype(np.uint8)
        rowptr = np.zeros(n + 1, dtype=np.int32)
        cap = min(n * n, 4 * M + n)
        colidx = np.zeros(max(cap, 1), dtype=np.int32)
        nnzb = L.wnfea_node_graph(n, M, 2, _lib.hptr(conn), _lib.hptr(node_free), _lib.hptr(rowptr),
                                  _lib.hptr(colidx), cap)
        if nnzb < 0:
            raise _lib.HipError(f"wnfea_node_graph failed ({nnzb})")
        colidx = colidx[:nnzb].copy()
        cptr = np.zeros(nnzb + 1, dtype=np.int64)
        codes = np.zeros(max(4 * M, 1), dtype=np.int32)
        nc = L.wnfea_contributions(n, M, 2, _lib.hptr(conn), _lib.hptr(rowptr), _lib.hptr(colidx),
                                   _lib.hptr(cptr), _lib.hptr(codes), 4 * M)
        if nc < 0:
            raise _lib.HipError(f"wnfea_contributions failed ({nc})")
        n2e_ptr = np.zeros(n + 1, dtype=np.int32)
        n2e = np.zeros(max(2 * M, 1), dtype=np.int32)
        _lib.check(L.wnfea_node_incidence(n, M, 2, _lib.hptr(conn), _lib.hptr(n2e_ptr), _lib.hptr(n2e)),
                   "wnfea_node_incidence")
        t = lambda a: torch.as_tensor(a, device=dev)  # noqa: E731
        self.d = {"nodes": t(nodes), "conn": t(conn), "props": t(props), "n2e_ptr": t(n2e_ptr), "n2e": t(n2e),
                  "cptr": t(cptr), "codes": t(codes[:max(nc, 1)].copy()), "fix": t(node_fix)}
        self.n_nodes, self.n_beams = n, M
        self.fe = torch.empty((max(M, 1), 12), dtype=torch.float64, device=dev)
        K = DeviceBSR(6, 6, n, n, rowptr, colidx, sigma)
        F = model.load_vector(6)
        F[fixed] = 0.0
        self.fixed = fixed
        self.fixed_t = torch.as_tensor(fixed, device=dev)
        diag = torch.empty(6 * n, dtype=torch.float64, device=dev)
        scale = torch.ones(6 * n, dtype=torch.float64, device=dev)
        self.system = SolidSystem(K=K, F=torch.as_tensor(F, device=dev), fixed_mask=fixed,
                                  fixed_dofs=model.fixed_dof_list(6), diag=diag, scale=scale, n_nodes=n,
                                  coords=nodes, mesh={}, dpn=6)
        self.amg = None
        self.cycle_tune = None
        _sync()
        self.timings = {"setup_sec": time.perf_counter() - t0}

    def close(self):
        if self.amg is not None:
            self.amg.close()
            self.amg = None

    # ------------------------------------------------------------ element kernels
    def internal_forces(self, U):
        """F_int(U) (device, 6 per node)."""
        import torch
        L = _lib.load()
        U = U.to(torch.float64).contiguous()
        F = torch.empty(6 * self.n_nodes, dtype=torch.float64, device=U.device)
        d = self.d
        _lib.check(L.wnfea_beam_forces(self.n_beams, self.n_nodes, _lib.dptr(d["conn"]), _lib.dptr(d["nodes"]),
                                       _lib.dptr(d["props"]), _lib.dptr(U), _lib.dptr(d["n2e_ptr"]),
                                       _lib.dptr(d["n2e"]), _lib.dptr(self.fe), _lib.dptr(F), 0, _lib.stream_of()),
                   "wnfea_beam_forces")
        return F

    def residual(self, U, lam: float):
        """R = F_int(U) - lam F_ext, R = U on fixed DOFs (residual.py:138-166)."""
        R = self.internal_forces(U) - lam * self.system.F
        R[self.fixed_t] = U[self.fixed_t]
        return R

    def _assemble(self, Ke, sym: int):
        from .solid_solver import refresh_scaling
        L = _lib.load()
        K, d = self.system.K, self.d
        _lib.check(L.wnfea_bsr_assemble_elem(K.handle, 2, sym, _lib.dptr(K.d_rowidx), _lib.dptr(K.d_colidx),
                                             _lib.dptr(d["cptr"]), _lib.dptr(d["codes"]), _lib.dptr(Ke),
                                             _lib.dptr(d["fix"]), None, _lib.stream_of()),
                   "wnfea_bsr_assemble_elem")
        refresh_scaling(self.system)

    def element_tangents(self, U):
        import torch
        L = _lib.load()
        Kt = torch.empty((max(self.n_beams, 1), 12, 12), dtype=torch.float64, device=U.device)
        d = self.d
        _lib.check(L.wnfea_beam_tangent(self.n_beams, _lib.dptr(d["conn"]), _lib.dptr(d["nodes"]),
                                        _lib.dptr(d["props"]), _lib.dptr(U.contiguous()), float(self.h_rel),
                                        _lib.dptr(Kt), _lib.stream_of()), "wnfea_beam_tangent")
        return Kt

    def assemble_tangent(self, U):
        """K = assembled symmetric part of the element tangents at U (Dirichlet
        rule applied), rescaled and re-equilibrated in place."""
        self._assemble(self.element_tangents(U), 1)

    def assemble_linear(self):
        """The reference's linear beam stiffness (assembler.py:13-148) in place."""
        import torch
        L = _lib.load()
        Ke = torch.empty((max(self.n_beams, 1), 12, 12), dtype=torch.float64, device=self.system.K.device)
        d = self.d
        _lib.check(L.wnfea_beam_ke_linear(self.n_beams, _lib.dptr(d["conn"]), _lib.dptr(d["nodes"]),
                                          _lib.dptr(d["props"]), _lib.dptr(Ke), _lib.stream_of()),
                   "wnfea_beam_ke_linear")
        self._assemble(Ke, 0)

    def build_amg(self, storage: tuple = ("fp32",), **kw):
        """SA-AMG of the current tangent, values kept in ``storage`` (FP32 by
        default; FP64 for an all-FP64 inner solve, FP32 + FP16 for the
        tri-precision switch).  The strength threshold
        defaults to 0.02: the 6-DOF lattice tangents converge fastest with
        weak filtering (40^3 lattice, 20 load steps: 3 120 CG iterations at
        theta 0.01-0.03, 3 460 at 0.08, 8 152 at 0.15), and the Chebyshev
        smoother's interval [lmax/ratio, lmax] defaults to ratio 20 (2 931
        iterations, 5 % faster than the solids' 30; 10-20 are alike, 60 takes
        3 652), and the prolongator smoothing factor to the textbook 4/3 (the
        solids' 1.6 takes 3 460 iterations).  Unless a launch shape is given,
        the cycle's SpMV launch shape is then tuned on whole cycles."""
        from .autotune import autotune_cycle
        from .solid_solver import build_amg
        self.close()
        kw.setdefault("theta", 0.02)
        kw.setdefault("ratio", 20.0)
        kw.setdefault("omega", 4.0 / 3.0)
        self.amg = build_amg(self.system, storage=tuple(storage), **kw)
        if self.amg.autotuned:
            self.cycle_tune = autotune_cycle(self.amg, "fp32" if "fp32" in storage else "fp64")
        return self.amg

    # ------------------------------------------------------------ Newton
    def solve_newton(self, n_load_steps: int = 20, max_newton_iter: int = 25, tol_rel: float = 1e-5,
                     tol_abs: float = 1e-5, inner_max_iter: int = 500, verbose: bool = False, amg_kw=None,
                     preconditioner: str = "amg", amg_precision: str = "fp32", tri_switch: float | None = None):
        """Load-stepped Newton (see module doc).  The inner solve is FP64 CG
        on the tangent, preconditioned by point Jacobi (``preconditioner=
        "jacobi"``) or by the SA-AMG V-cycle in ``amg_precision`` ("fp32" or
        "fp64"); with ``tri_switch`` the cycle runs with FP16-stored values
        until the inner relative residual is below tri_switch and with FP32
        values after that (the reference's tri-precision preconditioner,
        ``amg_preconditioner.py:538-584``).  Returns (U (device, 6 per
        node), stats)."""
        import torch
        from .solid_solver import pcg_device
        if preconditioner not in ("amg", "jacobi"):
            raise ValueError("preconditioner must be 'amg' or 'jacobi'")
        if amg_precision not in ("fp32", "fp64"):
            raise ValueError("amg_precision must be 'fp32' or 'fp64'")
        use_amg = preconditioner == "amg"
        storage = (amg_precision,) if tri_switch is None else ("fp32", "fp16")
        if tri_switch is not None:
            amg_precision = "fp32"
        t0 = time.perf_counter()
        dev = self.system.K.device
        U = torch.zeros(6 * self.n_nodes, dtype=torch.float64, device=dev)
        if use_amg and (self.amg is None or not set(storage) <= set(self.amg.storage)):
            self.assemble_tangent(U)
            self.build_amg(storage=storage, **dict(amg_kw or {}))
        stats = {"newton_iters": [], "pcg_iters": 0, "residual_evals": 0, "tangents": 0}
        for step, lam in enumerate(np.linspace(1.0 / n_load_steps, 1.0, n_load_steps), 1):
            R = self.residual(U, lam)
            stats["residual_evals"] += 1
            r0 = float(torch.linalg.norm(R))
            prev = r0
            its = 0
            if r0 < tol_abs:
                stats["newton_iters"].append(0)
                continue
            for it in range(1, max_newton_iter + 1):
                rn = float(torch.linalg.norm(R))
                if rn < tol_abs or rn / (r0 + 1e-20) < tol_rel:
                    break
                eta = min(0.1, 0.5 * (rn / (prev + 1e-20)) ** 2)
                inner_rel = max(eta * rn, 1e-10) / (rn + 1e-20)
                self.assemble_tangent(U)
                stats["tangents"] += 1
                if not use_amg:
                    dU, pits, _, _ = pcg_device(self.system, b=-R, rtol=inner_rel, max_iter=inner_max_iter,
                                                preconditioner="jacobi", precision="fp64")
                elif tri_switch is not None:
                    # FP16-stored cycles down to the switch, then FP32 cycles
                    # on the remaining residual
                    d16, p16, _, _ = pcg_device(self.system, b=-R, rtol=max(tri_switch, inner_rel),
                                                max_iter=inner_max_iter, preconditioner="amg", precision="fp64",
                                                amg=self.amg, amg_precision="fp16")
                    R16 = -R - self.system.K.spmv(d16)
                    rel16 = float(torch.linalg.norm(R16)) / (rn + 1e-300)
                    dU, pits = d16, p16
                    if rel16 > inner_rel:
                        d32, p32, _, _ = pcg_device(self.system, b=R16, rtol=inner_rel / rel16,
                                                    max_iter=inner_max_iter, preconditioner="amg",
                                                    precision="fp64", amg=self.amg, amg_precision="fp32")
                        dU, pits = d16 + d32, p16 + p32
                    stats["fp16_pcg_iters"] = stats.get("fp16_pcg_iters", 0) + p16
                else:
                    dU, pits, _, _ = pcg_device(self.system, b=-R, rtol=inner_rel, max_iter=inner_max_iter,
                                                preconditioner="amg", precision="fp64", amg=self.amg,
                                                amg_precision=amg_precision)
                stats["pcg_iters"] += pits
                prev = rn
                alpha = 1.0
                Ut = U + alpha * dU
                Rt = self.residual(Ut, lam)
                stats["residual_evals"] += 1
                tn = float(torch.linalg.norm(Rt))
                while tn >= rn and alpha > 0.0625:
                    alpha *= 0.5
                    Ut = U + alpha * dU
                    Rt = self.residual(Ut, lam)
                    stats["residual_evals"] += 1
                    tn = float(torch.linalg.norm(Rt))
                U, R = Ut, Rt
                its = it
                if verbose:
                    print(f"  step {step} newton {it}: |R| = {tn:.4e} (rel {tn / (r0 + 1e-20):.2e}) "
                          f"pcg {pits} alpha {alpha:.3f}")
            else:
                rn = float(torch.linalg.norm(R))
                if not (rn < tol_abs or rn / (r0 + 1e-20) < tol_rel):
                    raise NonLinearConvergenceError(
                        f"Newton solver failed to converge at load step {step} (lambda={lam:.3f}). "
                        f"Final residual: {rn:.4e}")
            stats["newton_iters"].append(its)
        torch.cuda.synchronize()
        stats["solve_sec"] = time.perf_counter() - t0
        stats["final_residual"] = float(torch.linalg.norm(self.residual(U, 1.0)))
        return U, stats


def compute_internal_forces_fast(model: FEAModel, U_full, F_int=None, device: str = "hip", block_size: int = 0,
                                 grid_size: int = 0):
    """Reference-compatible F_int (``fast_kernels.py:359-448``) of the beams
    and C3D10 solids of a model (6 DOF per node): U_full (N*6) numpy or
    device tensor -> numpy (or device tensor for device input); with F_int
    (numpy) the forces are added to it.  The device copies of the mesh are
    kept between calls; the launch-shape arguments are accepted and ignored."""
    from .residual import compute_internal_forces
    out = compute_internal_forces(model, U_full)
    if F_int is not None and isinstance(out, np.ndarray):
        F_int[:] += out
        return F_int
    return out


def solve_nonlinear_jfnk(model: FEAModel, n_load_steps: int = 5, max_newton_iter: int = 25, tol_rel: float = 1e-5,
                         tol_abs: float = 1e-5, use_amg: bool = True, use_mixed_precision: bool = False,
                         use_tri_precision: bool = False, switch_tol: float = 1e-2, verbose: bool = True):
    """Reference signature (``nonlinear_solver.py:159-170``); returns U_full
    (N*6, numpy) and stores it in ``model.displacements``.

    The flags select the inner solve as the reference's do
    (``nonlinear_solver.py:203-255``): ``use_amg`` the SA-AMG V-cycle (else
    point Jacobi, the reference's diagonal fallback); the cycle runs in FP64
    by default (the reference's FULL_FP64 preset), in FP32 with
    ``use_mixed_precision`` (MIXED_FP32), and with FP16-stored values above
    ``switch_tol`` and FP32 below it with ``use_tri_precision``
    (TRI_PRECISION).  The Newton residual, the Krylov vectors and the
    update stay FP64 in every mode."""
    errors = model.validate_for_solving()
    if errors:
        raise ValueError("Cannot solve model:\n  " + "\n  ".join(errors))
    if not (model.mesh_elements is not None and len(model.mesh_elements)) or \
            (getattr(model, "solid_elements", None) is not None and len(model.solid_elements)):
        raise ValueError("solve_nonlinear_jfnk runs the device Newton solver on beam-only models; mixed "
                         "beam-solid models go through compute_equilibrium_residual, MatrixFreeJFNKOperator "
                         "and pcg_solve")
    if use_tri_precision:
        amg_precision, tri = "fp32", float(switch_tol)
    else:
        amg_precision, tri = ("fp32" if use_mixed_precision else "fp64"), None
    bs = BeamSystem(model)
    try:
        U, _ = bs.solve_newton(n_load_steps=n_load_steps, max_newton_iter=max_newton_iter, tol_rel=tol_rel,
                               tol_abs=tol_abs, verbose=verbose, preconditioner="amg" if use_amg else "jacobi",
                               amg_precision=amg_precision, tri_switch=tri)
        out = U.cpu().numpy()
    finally:
        bs.close()
    model.displacements = out
    return out


def assemble_global_system_sparse(model: FEAModel, device: str = "auto"):
    """(K csr FP64, F_full) with 6 DOF per node, the reference's
    ``assembler.py:285-450``: linear beam stiffness and C3D10 solids (in the
    translational DOFs of their nodes) summed, then supports as zeroed
    rows/columns with a unit diagonal, F = 0 there, zeros eliminated.  The
    element matrices and their sums are computed on the device whatever
    ``device`` says (kept for API parity)."""
    import copy
    import scipy.sparse as sp
    has_beams = model.mesh_elements is not None and len(model.mesh_elements) > 0
    has_solids = getattr(model, "solid_elements", None) is not None and len(model.solid_elements) > 0
    n = len(model.mesh_nodes)
    fixed = model.fixed_dof_mask(6)
    F = model.load_vector(6)
    if has_beams and not has_solids:
        bs = BeamSystem(model)  # the device assembly applies the support rule itself
        bs.assemble_linear()
        K = bs.system.K.to_scipy_bsr().tocsr()
        bs.close()
    else:
        bare = copy.copy(model)
        bare.supports = []
        K = sp.csr_matrix((6 * n, 6 * n))
        if has_beams:
            bs = BeamSystem(bare)
            bs.assemble_linear()
            K = K + bs.system.K.to_scipy_bsr().tocsr()
            bs.close()
        if has_solids:
            from .solid_solver import assemble_c3d10_bsr
            s = assemble_c3d10_bsr(bare, apply_bcs=False)
            K3 = s.K.to_scipy_bsr().tocoo()
            s.K.close()
            m3 = (np.arange(3 * n) // 3) * 6 + np.arange(3 * n) % 3  # translational DOF in the 6-DOF layout
            K = K + sp.csr_matrix((K3.data, (m3[K3.row], m3[K3.col])), shape=(6 * n, 6 * n))
        K = sp.csr_matrix(K)
        if fixed.any():
            K.sort_indices()
            rows = np.repeat(np.arange(6 * n), np.diff(K.indptr))
            K.data[(fixed[rows] | fixed[K.indices]) & (rows != K.indices)] = 0.0
            K.data[fixed[rows] & (rows == K.indices)] = 1.0  # stored diagonals only, as the reference
    F[fixed] = 0.0
    K.eliminate_zeros()
    K.sort_indices()
    return K, F
