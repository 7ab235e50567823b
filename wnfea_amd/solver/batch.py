"""Many independent C3D10 meshes solved together on one GPU.

The meshes are concatenated into one block-diagonal system (node and element
offsets, supports and loads carried along), so every kernel of assembly,
SpMV, V-cycle and CG runs once for the whole batch instead of once per mesh:
small meshes that could not fill 256 CUs alone share every launch.  Nothing
couples the blocks: the node graph, the RCM order and the AMG aggregates stay
inside each mesh (the aggregation follows the strength graph, which has no
edges between meshes), and the coarsest dense inverse is block-diagonal.

Each mesh k must meet ||f_k - K_k u_k|| <= tol ||f_k||.  CG runs on the whole
system with the target tol * min_k ||f_k|| / ||f||, which implies every
per-mesh bound since ||r_k|| <= ||r||; the per-mesh residuals are then
verified in FP64 on the device.

The reference solves each model separately (``solid_solver.py:132-194``);
this is the batched path of the same solve.
"""
from __future__ import annotations

import copy
import time

import numpy as np

from ..model import FEAModel


def generate_box_batch(n_meshes: int = 64, seed: int = 1, dof_range=(2000, 60000), nx_range=(10, 60),
                       nyz_range=(3, 8), tip_load_total: float = -1000.0):
    """Seeded cantilever boxes (6-tet split, E = 210 GPa, nu = 0.3, x = 0
    clamped, tip load along -z split over the tip nodes) of uneven sizes:
    nx, ny = nz drawn uniformly until 3 (2nx+1)(2ny+1)^2 DOF fall in
    dof_range.  Box length nx/40 m and side ny/40 m (the 40x4x4 box's cell)."""
    from ..mesh.solid_generator import generate_c3d10_structured_block
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n_meshes:
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(nyz_range[0], nyz_range[1] + 1))
        dof = 3 * (2 * nx + 1) * (2 * ny + 1) ** 2
        if not (dof_range[0] <= dof <= dof_range[1]):
            continue
        model, meta = generate_c3d10_structured_block(nx / 40.0, ny / 40.0, ny / 40.0, nx, ny, ny,
                                                      tip_load_total=tip_load_total, tets_per_hex=6,
                                                      load_direction="z")
        out.append((model, meta))
    return out


def merge_models(models: list[FEAModel]) -> tuple[FEAModel, np.ndarray]:
    """One model holding all meshes; node_offsets[k] is the first node of mesh k
    (length n_meshes + 1)."""
    offs = np.zeros(len(models) + 1, dtype=np.int64)
    for k, m in enumerate(models):
        offs[k + 1] = offs[k] + len(m.mesh_nodes)
    merged = FEAModel()
    merged.mesh_nodes = np.concatenate([m.mesh_nodes for m in models]).astype(np.float64)
    merged.solid_elements = np.concatenate(
        [m.solid_elements.astype(np.int64) + offs[k] for k, m in enumerate(models)]).astype(np.int32)
    Es, nus = [], []
    for m in models:
        E, nu = m.element_moduli()
        Es.append(E)
        nus.append(nu)
    merged.solid_E = np.concatenate(Es)
    merged.solid_nu = np.concatenate(nus)
    merged.materials = dict(models[0].materials)
    for k, m in enumerate(models):
        for s in m.supports:
            c = copy.copy(s)
            c.node_id = int(s.node_id + offs[k])
            merged.supports.append(c)
        for ld in m.loads:
            c = copy.copy(ld)
            c.node_id = int(ld.node_id + offs[k])
            merged.loads.append(c)
    return merged, offs


class BatchedSolids:
    """Assemble, precondition and solve a list of C3D10 models as one system."""

    def __init__(self, models: list[FEAModel], amg_precision: str = "fp32", rcm: bool = True, amg_kw=None,
                 operator: str = "matrix_free"):
        from ..mesh.rcm import apply_rcm_to_model
        from .solid_solver import MatrixFreeOperator, assemble_c3d10_bsr, build_amg
        t0 = time.perf_counter()
        merged, offs = merge_models(models)
        self.n2o = None
        if rcm:
            merged, self.n2o, _ = apply_rcm_to_model(merged)
        self.merged, self.offsets = merged, offs
        # mesh id of every (RCM-ordered) node, for per-mesh norms
        owner = np.repeat(np.arange(len(models)), np.diff(offs))
        self.owner_host = owner[self.n2o] if self.n2o is not None else owner
        self.system = assemble_c3d10_bsr(merged, want_fp64_equilibrated=(amg_precision == "fp64"))
        t1 = time.perf_counter()
        storage = ("fp64",) if amg_precision == "fp64" else ("fp32",)
        self.amg = build_amg(self.system, storage=storage, **dict(amg_kw or {}))
        self.amg_precision = amg_precision
        # CG and the finest smoothing steps on the matrix-free operator (default)
        self.op = None
        if operator == "matrix_free":
            self.op = MatrixFreeOperator(self.system)
            self.op.attach_smoother(self.amg)
        elif operator != "assembled":
            raise ValueError(f"unknown operator {operator!r}")
        self.timings = {"assembly_sec": t1 - t0, "amg_setup_sec": time.perf_counter() - t1}
        import torch
        dev = self.system.K.device
        self._owner = torch.as_tensor(np.repeat(self.owner_host, 3), device=dev)
        self.n_meshes = len(models)
        f2 = torch.zeros(self.n_meshes, dtype=torch.float64, device=dev)
        f2.index_add_(0, self._owner, self.system.F * self.system.F)
        self.f_norms = torch.sqrt(f2)

    def close(self):
        self.amg.close()
        if self.op is not None:
            self.op.close()
            self.op = None

    def per_mesh_residuals(self, u):
        """||f_k - K_k u_k|| / ||f_k|| for every mesh (device FP64 residual)."""
        import torch
        r = self.system.F - self.system.K.spmv(u)
        r2 = torch.zeros(self.n_meshes, dtype=torch.float64, device=u.device)
        r2.index_add_(0, self._owner, r * r)
        return (torch.sqrt(r2) / torch.where(self.f_norms > 0, self.f_norms, torch.ones_like(self.f_norms)))

    def solve(self, tol: float = 1e-8):
        """All meshes to tol (FP64-verified per mesh).  Returns (u in the merged
        RCM order, iterations, per-mesh relative residuals (host), seconds)."""
        import torch
        from .solid_solver import solve_mixed
        t0 = time.perf_counter()
        fmin = float(self.f_norms[self.f_norms > 0].min()) if bool((self.f_norms > 0).any()) else 0.0
        fall = float(torch.linalg.norm(self.system.F))
        tol_global = tol * fmin / fall if fall > 0 else tol
        u, iters, runs, rel, _ = solve_mixed(self.system, self.amg, tol=tol_global,
                                             amg_precision=self.amg_precision, max_restarts=4, operator=self.op)
        per = self.per_mesh_residuals(u)
        torch.cuda.synchronize()
        return u, iters, per.cpu().numpy(), time.perf_counter() - t0

    def split(self, u) -> list[np.ndarray]:
        """Per-mesh displacement vectors (3 per node, each mesh's own node order)."""
        u_host = u.detach().cpu().numpy().reshape(-1, 3)
        if self.n2o is not None:
            full = np.empty_like(u_host)
            full[self.n2o] = u_host
            u_host = full
        return [u_host[self.offsets[k]:self.offsets[k + 1]].ravel().copy() for k in range(self.n_meshes)]
