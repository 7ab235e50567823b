"""C3D10 assembly and solvers on the GPU (drop-in for ``wnfea/solver/solid_solver.py``).

Reference functions kept with their signatures and return conventions:

* ``assemble_c3d10_sparse_3dof(model, apply_bcs=True) -> (K csr, F, fixed_dofs)``
  (``solid_solver.py:26-129``): element stiffness, global assembly and the
  Dirichlet rule run on the device; the CSR is exported for drop-in callers.
* ``solve_c3d10_system(K, F, method, rtol, max_iter, preconditioner, precision)``
  and ``solve_c3d10_pcg(...)`` (``:132-308``): device PCG (Jacobi, block
  Jacobi, none or AMG), FP64 or FP32-inner iterative refinement.
* ``compute_c3d10_stress_field(model, u_3dof)`` (``:311-365``).

The device-resident path used by the benchmarks is ``assemble_c3d10_bsr`` ->
``SolidSystem`` (K on the device as sliced-ELL 3x3 BSR) -> ``solve_system``.
There is no CPU fallback anywhere in this module.
"""
from __future__ import annotations

import os
import time
import weakref
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from .. import _lib
from ..model import FEAModel
from .bsr_matrix import DeviceBSR

_NPE = 10


def _sync():
    import torch
    torch.cuda.synchronize()


@dataclass
class SolidSystem:
    """Assembled 3-DOF system resident on the device."""
    K: DeviceBSR                      # FP64 K (BCs applied), FP32 S K S, optional FP64 S K S
    F: object                         # torch f64 (n_dofs,)
    fixed_mask: np.ndarray            # bool (n_dofs,)
    fixed_dofs: list
    diag: object                      # torch f64 diag(K)
    scale: object                     # torch f64 1/sqrt|diag| (1 where |diag| <= 1e-12)
    n_nodes: int
    coords: np.ndarray | None = None
    mesh: dict = field(default_factory=dict)   # device mesh arrays for element kernels
    timings: dict = field(default_factory=dict)
    dpn: int = 3                      # DOF per node (block size): 3 solids, 6 beams

    @property
    def n_dofs(self) -> int:
        return self.dpn * self.n_nodes

    def equilibrated_host(self) -> sp.bsr_matrix:
        """S K S as a host scipy BSR (FP64), the input of the AMG setup."""
        import torch
        vals = self.K.get_values()
        s = self.scale
        rows = self.K.d_rowidx.long()
        cols = self.K.d_colidx.long()
        b = self.dpn
        sr = s.view(-1, b)[rows]  # (nnzb, b)
        sc = s.view(-1, b)[cols]
        v = vals.view(-1, b, b) * sr[:, :, None] * sc[:, None, :]
        out = sp.bsr_matrix((v.cpu().numpy(), self.K.colidx, self.K.rowptr), shape=self.K.shape)
        del vals, sr, sc, v
        torch.cuda.empty_cache()
        return out


def _diag_block_index(rowptr: np.ndarray, colidx: np.ndarray) -> np.ndarray:
    n = len(rowptr) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr))
    pos = np.flatnonzero(rows == colidx)
    out = np.full(n, -1, dtype=np.int64)
    out[rows[pos]] = pos
    return out


def _finish_system(K: DeviceBSR, F_host: np.ndarray, fixed_mask: np.ndarray, fixed_dofs: list, n_nodes: int,
                   want_fp64_equilibrated: bool, coords=None, mesh=None, timings=None) -> SolidSystem:
    import torch
    L = _lib.load()
    dev = K.device
    stream = _lib.stream_of()
    dpn = K.br
    dblk = torch.as_tensor(_diag_block_index(K.rowptr, K.colidx), device=dev)
    diag = torch.empty(dpn * n_nodes, dtype=torch.float64, device=dev)
    _lib.check(L.wnfea_bsr_diag(K.handle, _lib.dptr(dblk), _lib.dptr(diag), stream), "wnfea_bsr_diag")
    ad = diag.abs()
    scale = torch.where(ad > 1e-12, 1.0 / torch.sqrt(torch.where(ad > 1e-12, ad, torch.ones_like(ad))),
                        torch.ones_like(ad)).contiguous()
    _lib.check(L.wnfea_bsr_equilibrate(K.handle, _lib.dptr(K.d_rowidx), _lib.dptr(K.d_colidx), _lib.dptr(scale),
                                       int(want_fp64_equilibrated), 0, stream), "wnfea_bsr_equilibrate")
    F = torch.as_tensor(np.ascontiguousarray(F_host, dtype=np.float64), device=dev)
    return SolidSystem(K=K, F=F, fixed_mask=fixed_mask, fixed_dofs=fixed_dofs, diag=diag, scale=scale,
                       n_nodes=n_nodes, coords=coords, mesh=mesh or {}, timings=timings or {}, dpn=dpn)


def mesh_to_device(model: FEAModel, dev=None) -> dict:
    """Device copies of the arrays the element kernels read."""
    import torch
    dev = dev or _lib.require_gpu()
    L = _lib.load()
    nodes = np.ascontiguousarray(model.mesh_nodes, dtype=np.float64)
    conn = np.ascontiguousarray(model.solid_elements, dtype=np.int32)
    if conn.ndim != 2 or conn.shape[1] != _NPE:
        raise ValueError("solid_elements must have shape (M, 10)")
    n, M = len(nodes), len(conn)
    if M and (conn.min() < 0 or conn.max() >= n):
        raise ValueError("solid_elements reference a node outside mesh_nodes")
    E, nu = model.element_moduli()
    n2e_ptr = np.zeros(n + 1, dtype=np.int32)
    n2e = np.zeros(max(M * _NPE, 1), dtype=np.int32)
    _lib.check(L.wnfea_node_incidence(n, M, _NPE, _lib.hptr(conn), _lib.hptr(n2e_ptr), _lib.hptr(n2e)),
               "wnfea_node_incidence")
    t = lambda a: torch.as_tensor(a, device=dev)  # noqa: E731
    return {"nodes": t(nodes), "conn": t(conn), "E": t(np.ascontiguousarray(E)), "nu": t(np.ascontiguousarray(nu)),
            "n2e_ptr": t(n2e_ptr), "n2e": t(n2e), "n_nodes": n, "n_elements": M, "conn_host": conn,
            "E_host": E, "nu_host": nu}


def assemble_c3d10_bsr(model: FEAModel, apply_bcs: bool = True, sigma: int = 16384,
                       want_fp64_equilibrated: bool = False, mesh: dict | None = None,
                       keep_plan: bool = False) -> SolidSystem:
    """Assemble K (3x3 BSR on the device), F and the Dirichlet data of a C3D10 model.

    keep_plan keeps the device contribution lists so that new element moduli
    can be assembled into the same pattern (``reassemble_c3d10_values``)."""
    import torch
    t0 = time.perf_counter()
    L = _lib.load()
    dev = _lib.require_gpu()
    mesh = mesh or mesh_to_device(model, dev)
    n, M = mesh["n_nodes"], mesh["n_elements"]
    conn = mesh["conn_host"]
    fixed_mask = model.fixed_dof_mask(3) if apply_bcs else np.zeros(3 * n, dtype=bool)
    fixed_dofs = model.fixed_dof_list(3) if apply_bcs else []
    node_fix = (fixed_mask.reshape(n, 3) * np.array([1, 2, 4])).sum(axis=1).astype(np.uint8)
    node_free = (node_fix != 7).astype(np.uint8)
    # block pattern with the Dirichlet rule
    rowptr = np.zeros(n + 1, dtype=np.int32)
    cap = min(n * n, M * _NPE * _NPE + n)
    colidx = np.zeros(max(cap, 1), dtype=np.int32)
    nnzb = L.wnfea_node_graph(n, M, _NPE, _lib.hptr(conn), _lib.hptr(node_free) if apply_bcs else None,
                              _lib.hptr(rowptr), _lib.hptr(colidx), cap)
    if nnzb < 0:
        raise _lib.HipError(f"wnfea_node_graph failed ({nnzb})")
    colidx = colidx[:nnzb].copy()
    cptr = np.zeros(nnzb + 1, dtype=np.int64)
    codes = np.zeros(max(M * _NPE * _NPE, 1), dtype=np.int32)
    nc = L.wnfea_contributions(n, M, _NPE, _lib.hptr(conn), _lib.hptr(rowptr), _lib.hptr(colidx), _lib.hptr(cptr),
                               _lib.hptr(codes), M * _NPE * _NPE)
    if nc < 0:
        raise _lib.HipError(f"wnfea_contributions failed ({nc})")
    t_plan = time.perf_counter() - t0
    K = DeviceBSR(3, 3, n, n, rowptr, colidx, sigma)
    t_layout = time.perf_counter() - t0 - t_plan
    stream = _lib.stream_of()
    Ke = torch.empty((M, 30, 30), dtype=torch.float64, device=dev)
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.check(L.wnfea_c3d10_ke(M, _lib.dptr(mesh["conn"]), _lib.dptr(mesh["nodes"]), _lib.dptr(mesh["E"]),
                                _lib.dptr(mesh["nu"]), _lib.dptr(Ke), _lib.dptr(bad), stream), "wnfea_c3d10_ke")
    d_cptr = torch.as_tensor(cptr, device=dev)
    d_codes = torch.as_tensor(codes[:nc].copy() if nc else codes[:1].copy(), device=dev)
    d_fix = torch.as_tensor(node_fix, device=dev)
    _lib.check(L.wnfea_bsr_assemble(K.handle, _lib.dptr(K.d_rowidx), _lib.dptr(K.d_colidx), _lib.dptr(d_cptr),
                                    _lib.dptr(d_codes), _lib.dptr(Ke), _lib.dptr(d_fix) if apply_bcs else None,
                                    None, stream), "wnfea_bsr_assemble")
    if int(bad.item()):
        raise ValueError(f"Non-positive Jacobian determinant in C3D10 element {int(bad.item()) - 1}. "
                         "Element is inverted, degenerate, or has bad node numbering.")
    F = model.load_vector(3)
    if apply_bcs:
        F[fixed_mask] = 0.0
    sys_ = _finish_system(K, F, fixed_mask, fixed_dofs, n, want_fp64_equilibrated,
                          coords=np.asarray(model.mesh_nodes, dtype=np.float64), mesh=mesh)
    _sync()
    if keep_plan:
        mesh["plan"] = {"cptr": d_cptr, "codes": d_codes, "fix": d_fix if apply_bcs else None,
                        "want64e": bool(want_fp64_equilibrated)}
    del Ke, d_cptr, d_codes
    torch.cuda.empty_cache()
    sys_.timings = {"plan_sec": t_plan, "layout_sec": t_layout, "assemble_total_sec": time.perf_counter() - t0}
    return sys_


def reassemble_c3d10_values(system: SolidSystem, E=None, nu=None) -> SolidSystem:
    """Assemble new per-element moduli (device tensors or arrays of length M)
    into the resident K of a system built with keep_plan=True: element
    stiffnesses, the deterministic gather, the diagonal, the scaling and the
    equilibrated copies are recomputed in place; the pattern is unchanged."""
    import torch
    L = _lib.load()
    mesh = system.mesh
    plan = mesh.get("plan")
    if plan is None:
        raise ValueError("system was assembled without keep_plan=True")
    dev = system.K.device
    M = mesh["n_elements"]
    for key, val in (("E", E), ("nu", nu)):
        if val is not None:
            t = torch.as_tensor(val, dtype=torch.float64, device=dev).reshape(-1)
            if t.numel() != M:
                raise ValueError(f"{key} must have one value per element ({M})")
            # in place: matrix-free operators of this system read the same buffer
            mesh[key].copy_(t)
    stream = _lib.stream_of()
    Ke = torch.empty((M, 30, 30), dtype=torch.float64, device=dev)
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.check(L.wnfea_c3d10_ke(M, _lib.dptr(mesh["conn"]), _lib.dptr(mesh["nodes"]), _lib.dptr(mesh["E"]),
                                _lib.dptr(mesh["nu"]), _lib.dptr(Ke), _lib.dptr(bad), stream), "wnfea_c3d10_ke")
    K = system.K
    _lib.check(L.wnfea_bsr_assemble(K.handle, _lib.dptr(K.d_rowidx), _lib.dptr(K.d_colidx), _lib.dptr(plan["cptr"]),
                                    _lib.dptr(plan["codes"]), _lib.dptr(Ke),
                                    _lib.dptr(plan["fix"]) if plan["fix"] is not None else None, None, stream),
               "wnfea_bsr_assemble")
    if int(bad.item()):
        raise ValueError(f"Non-positive Jacobian determinant in C3D10 element {int(bad.item()) - 1}.")
    del Ke
    refresh_scaling(system)
    return system


def refresh_scaling(system: SolidSystem) -> SolidSystem:
    """After new FP64 values of K (same pattern): diag(K), S = diag(1/sqrt|K_ii|)
    and the equilibrated copies (FP32, and FP64/FP16 where they exist)."""
    import torch
    L = _lib.load()
    K = system.K
    stream = _lib.stream_of()
    dblk = system.mesh.get("_diag_block")
    if dblk is None:
        dblk = torch.as_tensor(_diag_block_index(K.rowptr, K.colidx), device=K.device)
        system.mesh["_diag_block"] = dblk
    _lib.check(L.wnfea_bsr_diag(K.handle, _lib.dptr(dblk), _lib.dptr(system.diag), stream), "wnfea_bsr_diag")
    ad = system.diag.abs()
    system.scale.copy_(torch.where(ad > 1e-12, 1.0 / torch.sqrt(torch.where(ad > 1e-12, ad, torch.ones_like(ad))),
                                   torch.ones_like(ad)))
    info = K.info()
    _lib.check(L.wnfea_bsr_equilibrate(K.handle, _lib.dptr(K.d_rowidx), _lib.dptr(K.d_colidx),
                                       _lib.dptr(system.scale), int(bool(info["has_fp64_equilibrated"])),
                                       int(bool(info.get("has_fp16", False))), stream), "wnfea_bsr_equilibrate")
    _sync()
    return system


def system_from_csr(K_csr, F, fixed_mask: np.ndarray | None = None, coords=None,
                    want_fp64_equilibrated: bool = False, sigma: int = 4096, dpn: int = 3) -> SolidSystem:
    """Device system from a host scalar CSR (``dpn`` = 3 or 6 DOF per node),
    e.g. a drop-in caller's K."""
    import torch
    from .bsr_matrix import csr_to_bsr
    if dpn not in (3, 6):
        raise ValueError("dpn must be 3 or 6")
    Kb = csr_to_bsr(K_csr, dpn, np.float64)
    n = Kb.n_nodes
    K = DeviceBSR(dpn, dpn, n, n, Kb.block_row_ptr, Kb.block_col_idx, sigma)
    K.set_values(torch.as_tensor(Kb.values, device=K.device), which=1)
    if fixed_mask is None:
        fixed_mask = np.zeros(dpn * n, dtype=bool)
    return _finish_system(K, np.asarray(F, dtype=np.float64), fixed_mask, list(np.flatnonzero(fixed_mask)), n,
                          want_fp64_equilibrated, coords=coords)


def assemble_c3d10_sparse_3dof(model: FEAModel, apply_bcs: bool = True) -> tuple[sp.csr_matrix, np.ndarray, list]:
    """Reference-compatible assembly: (K csr FP64, F, fixed_dofs); computed on the device."""
    s = assemble_c3d10_bsr(model, apply_bcs=apply_bcs)
    K = s.K.to_scipy_bsr().tocsr()
    if apply_bcs and s.fixed_dofs:
        K.eliminate_zeros()
    K.sort_indices()
    F = s.F.cpu().numpy()
    s.K.close()
    return K, F, list(s.fixed_dofs)


# ------------------------------------------------------------------ solvers
class _Solver:
    """Process-wide solver workspace (Krylov vectors, reduction buffers and the
    replayable iteration graph).  WNFEA_GRAPHS=0 turns graph replay off."""
    _h = None

    @classmethod
    def get(cls):
        if cls._h is None:
            L = _lib.load()
            cls._h = L.wnfea_solver_create()
            if not cls._h:
                raise _lib.HipError("wnfea_solver_create failed")
            _lib.check(L.wnfea_solver_set_graphs(cls._h, int(os.environ.get("WNFEA_GRAPHS", "0") != "0")),
                       "wnfea_solver_set_graphs")
        return cls._h


def set_graphs(enable: bool) -> None:
    """Run full chunks of PCG iterations as a captured HIP graph, or launch
    every kernel from the host (default: on the I-beam solve host launches
    are about 3 % faster than graph replay, the queue never runs dry)."""
    L = _lib.load()
    _lib.check(L.wnfea_solver_set_graphs(_Solver.get(), int(bool(enable))), "wnfea_solver_set_graphs")


def _as_system(K, F, want_fp64_equilibrated=False) -> SolidSystem:
    if isinstance(K, SolidSystem):
        return K
    return system_from_csr(K, F, want_fp64_equilibrated=want_fp64_equilibrated)


_VT = {"fp64": 0, "fp32": 1, "fp16": 2, "fp16_fine": 3, "fp16_coarse": 4, "fp16_smooth": 5}


def pcg_device(system: SolidSystem, b=None, rtol: float = 1e-8, max_iter: int = 2000, preconditioner: str = "jacobi",
               precision: str = "fp64", amg=None, amg_precision: str = "fp64", check_every: int = 16):
    """One device PCG on the system.

    precision (the Krylov operator and vectors): "fp64" (K, FP64 vectors),
    "fp64_equilibrated" (S K S), "fp32" (S K S in FP32, FP32 vectors, FP64
    dots).  With preconditioner="amg" on "fp64", the V-cycle runs on S K S as
    z = S M(S r) in ``amg_precision`` ("fp64", "fp32", "fp16" storage with
    FP32 arithmetic, or "fp16_fine": FP16 storage on the finest level and FP32
    below): FP64 CG with a lower-precision AMG preconditioner.
    Returns (x (device), iterations, recursive relative residual, seconds)."""
    import torch
    L = _lib.load()
    dev = system.K.device
    n = system.n_dofs
    prec = {"fp64": 0, "fp32": 1, "fp64_equilibrated": 2}[precision]
    dt = torch.float32 if prec == 1 else torch.float64
    b = system.F if b is None else b
    if prec != 0:
        b = (b * system.scale).to(dt)  # equilibrated right-hand side
    pkind = {"none": 0, "jacobi": 1, "amg": 2}[preconditioner]
    dinv = None
    if pkind == 1:
        dinv = (1.0 / system.diag).to(dt) if prec == 0 else torch.ones(n, dtype=dt, device=dev)
    if pkind == 2 and amg is None:
        raise ValueError("preconditioner='amg' needs an amg handle")
    vt = _VT[amg_precision] if pkind == 2 else 0
    if pkind == 2 and prec == 1 and vt == 0:
        raise ValueError("an FP32 Krylov loop takes an FP32 or FP16 V-cycle")
    scale = system.scale if (pkind == 2 and prec == 0) else None
    x = torch.empty(n, dtype=dt, device=dev)
    it = _lib.ctypes.c_int(0)
    rel = _lib.ctypes.c_double(0.0)
    _sync()
    t0 = time.perf_counter()
    _lib.check(L.wnfea_pcg(_Solver.get(), system.K.handle, prec, pkind, _lib.dptr(dinv) if dinv is not None else None,
                           amg.handle if amg is not None else None, vt, _lib.dptr(scale) if scale is not None else None,
                           _lib.dptr(b), _lib.dptr(x), float(rtol), int(max_iter), int(check_every),
                           _lib.ctypes.byref(it), _lib.ctypes.byref(rel), _lib.stream_of()), "wnfea_pcg")
    _sync()
    dt_s = time.perf_counter() - t0
    if prec != 0:
        x = x.to(torch.float64) * system.scale
    return x, it.value, rel.value, dt_s


class MatrixFreeOperator:
    """K of a C3D10 system applied without the assembled matrix: per Gauss
    point inverse Jacobians and weights (320 B per element, "partial
    assembly") and the element contraction B^T D B on the fly, with the
    assembled matrix's Dirichlet rule (fixed DOF: unit diagonal, no
    coupling).  The reference's matrix-free f_int kernel used as a solve
    operator (``hip_kernels.cpp:283-425``): C1 needs 67 MB of quadrature data
    instead of 635 MB of FP64 blocks."""

    def __init__(self, system: SolidSystem):
        import torch
        if system.dpn != 3 or not system.mesh:
            raise ValueError("the matrix-free operator needs a C3D10 system with its device mesh")
        self.lib = _lib.load()
        self.system = system
        m = system.mesh
        bits = np.asarray(system.fixed_mask, dtype=np.uint8).reshape(-1, 3) * np.array([1, 2, 4], dtype=np.uint8)
        self.fix = torch.as_tensor(bits.sum(axis=1).astype(np.uint8), device=m["nodes"].device)
        bad = torch.zeros(1, dtype=torch.int32, device=m["nodes"].device)
        h = _lib.ctypes.c_void_p()
        _lib.check(self.lib.wnfea_mf_create(m["n_elements"], m["n_nodes"], _lib.dptr(m["conn"]), _lib.dptr(m["nodes"]),
                                            _lib.dptr(m["E"]), _lib.dptr(m["nu"]), _lib.dptr(self.fix),
                                            _lib.dptr(m["n2e_ptr"]), _lib.dptr(m["n2e"]), _lib.dptr(bad),
                                            _lib.ctypes.byref(h), _lib.stream_of()), "wnfea_mf_create")
        self.handle = h.value
        # the library borrows these device buffers: keep them alive with the handle
        self._borrowed = tuple(m[k] for k in ("conn", "nodes", "E", "nu", "n2e_ptr", "n2e"))
        self._attached = weakref.WeakSet()  # hierarchies smoothing on this operator
        if int(bad.item()):
            self.close()
            raise ValueError(f"Non-positive Jacobian determinant in C3D10 element {int(bad.item()) - 1}")
        self.n = 3 * m["n_nodes"]

    @property
    def affine(self) -> bool:
        """True when every element is affine and one inverse Jacobian per
        element is stored (else one per Gauss point)."""
        r = self.lib.wnfea_mf_affine(self.handle)
        if r < 0:
            _lib.check(r, "wnfea_mf_affine")
        return r == 1

    def attach_smoother(self, amg):
        """Make ``amg``'s finest-level smoothing and defect products use this
        operator (SmoothedAggregationAMG.attach_matrix_free)."""
        amg.attach_matrix_free(self)
        return amg

    def set_patch(self, elems: int):
        """Rebuild the patch layout of the products with ``elems`` elements
        per patch (a multiple of 64, at most 256), or drop it (0: per-element
        contributions and a node gather).  Affine meshes only; the layout is
        built with 64 at construction when it fits.  Hierarchies smoothing
        on this operator are attached again (their FP32 copies follow the
        layout)."""
        _lib.check(self.lib.wnfea_mf_set_patch(self.handle, int(elems), _lib.stream_of()), "wnfea_mf_set_patch")
        for amg in list(self._attached):
            if getattr(amg, "handle", None):
                amg.attach_matrix_free(self, getattr(amg, "_mf_scale", None))

    @property
    def patch_info(self) -> dict:
        """Elements per patch (0: no patch layout), patches, the most local
        nodes of one patch, the local nodes (= partial sums) of all and the
        per-patch stride of the local-node arrays."""
        out = (_lib.ctypes.c_longlong * 6)()
        _lib.check(self.lib.wnfea_mf_info(self.handle, out), "wnfea_mf_info")
        return {"elems": int(out[0]), "patches": int(out[1]), "lmax": int(out[2]), "local_nodes": int(out[3]),
                "stride": int(out[5])}

    def apply(self, u, y=None):
        import torch
        y = torch.empty_like(u) if y is None else y
        _lib.check(self.lib.wnfea_mf_apply(self.handle, _lib.dptr(u), _lib.dptr(y), _lib.stream_of()),
                   "wnfea_mf_apply")
        return y

    def residual(self, u, f=None) -> float:
        """||f - K u|| / ||f|| (f defaults to the system's F)."""
        import torch
        f = self.system.F if f is None else f
        r = torch.empty_like(f)
        out = (_lib.ctypes.c_double * 2)()
        _lib.check(self.lib.wnfea_mf_residual(_Solver.get(), self.handle, _lib.dptr(u), _lib.dptr(f), _lib.dptr(r),
                                              out, _lib.stream_of()), "wnfea_mf_residual")
        return out[0] / out[1] if out[1] > 0 else 0.0

    def close(self):
        for amg in list(getattr(self, "_attached", ())):
            if getattr(amg, "handle", None):
                amg.attach_matrix_free(None)
        if getattr(self, "handle", None):
            self.lib.wnfea_mf_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_mixed(system: SolidSystem, amg, tol: float = 1e-8, amg_precision: str = "fp32", max_iter: int = 2000,
                check_every: int = 16, margin: float = 0.5, max_restarts: int = 3,
                operator: MatrixFreeOperator | None = None):
    """Mixed-precision solve of the resident system to ||F - K u|| <= tol ||F||.

    FP64 CG on K (FP64 SpMV, FP64 alpha/beta, FP64 solution update) with the
    SA-AMG V-cycle in ``amg_precision`` on S K S.  The stopping test of CG
    uses its recursive residual, so the run targets margin*tol and the FP64
    true residual is verified afterwards; if it is above tol (recursion drift)
    CG is restarted on the true residual (a refinement step).  Restarts stop
    early once a step no longer reduces the true residual by half: the FP64
    residual has then reached its round-off floor.  With ``operator`` (a
    MatrixFreeOperator of the system) CG's products, the true residual and
    the restarts use the matrix-free K; the V-cycle is unchanged.
    Returns (u, iterations, CG runs, ||F-Ku||/||F||, seconds)."""
    t0 = time.perf_counter()
    u = None
    iters = 0
    steps = 0
    b = system.F
    rel = float("inf")
    for steps in range(1, max_restarts + 2):
        rt = margin * tol * (1.0 if u is None else float(system.F.norm() / b.norm()))
        if operator is None:
            x, it, _, _ = pcg_device(system, b=b, rtol=rt, max_iter=max_iter, preconditioner="amg", precision="fp64",
                                     amg=amg, amg_precision=amg_precision, check_every=check_every)
        else:
            x, it = _pcg_mf(system, operator, amg, b, rt, max_iter, amg_precision, check_every)
        iters += it
        u = x if u is None else u + x
        prev, rel = rel, (true_residual(system, u) if operator is None else operator.residual(u))
        if rel <= tol or steps > max_restarts or rel > 0.5 * prev:
            break
        b = system.F - (system.K.spmv(u) if operator is None else operator.apply(u))
    _sync()
    return u, iters, steps, rel, time.perf_counter() - t0


def _pcg_mf(system: SolidSystem, operator: MatrixFreeOperator, amg, b, rtol: float, max_iter: int,
            amg_precision: str, check_every: int):
    """FP64 AMG-PCG from x = 0 on the matrix-free operator; (x, iterations)."""
    import torch
    L = _lib.load()
    x = torch.empty_like(b)
    it = _lib.ctypes.c_int(0)
    rel = _lib.ctypes.c_double(0.0)
    _lib.check(L.wnfea_pcg_mf(_Solver.get(), operator.handle, 2, None, amg.handle, amg.VT_CODE[amg_precision],
                              _lib.dptr(system.scale), _lib.dptr(b), _lib.dptr(x), float(rtol), int(max_iter),
                              int(check_every), _lib.ctypes.byref(it), _lib.ctypes.byref(rel), _lib.stream_of()),
               "wnfea_pcg_mf")
    return x, it.value


def solve_multi(system: SolidSystem, amg, B, tol: float = 1e-8, group: int = 8, max_iter: int = 2000,
                check_every: int = 16, margin: float = 0.5, max_restarts: int = 6):
    """Many load cases at once: every column of B (device FP64, n_dofs x k) to
    ||b_k - K u_k|| <= tol ||b_k|| (FP64-verified per column).

    Columns go through the device solver in groups of `group` (4 or 8): one
    SpMM per iteration reads K once for the whole group, and the V-cycle runs
    on all columns of the group at once (FP64 CG per column, FP32 SA-AMG on
    S K S).  A column whose true residual misses tol after the run is
    restarted on its true residual, until it meets tol or a restart no longer
    halves it (the FP64 residual floor).  Returns (U (n x k), iterations per
    column, relative residual per column, seconds)."""
    import torch
    L = _lib.load()
    if group not in (4, 8):
        raise ValueError("group must be 4 or 8")
    B = torch.as_tensor(B, dtype=torch.float64, device=system.K.device)
    if B.dim() == 1:
        B = B[:, None]
    n, k = B.shape
    if n != system.n_dofs:
        raise ValueError("B must have n_dofs rows")
    t0 = time.perf_counter()
    U = torch.zeros_like(B)
    iters = np.zeros(k, dtype=np.int64)
    rels = np.zeros(k)
    its = (_lib.ctypes.c_int * group)()
    rel = (_lib.ctypes.c_double * group)()
    h = _Solver.get()
    for g0 in range(0, k, group):
        m = min(group, k - g0)
        Bg = torch.zeros((n, group), dtype=torch.float64, device=B.device)
        Bg[:, :m] = B[:, g0:g0 + m]
        Xg = torch.zeros_like(Bg)
        bn = torch.linalg.norm(Bg, dim=0)
        rhs = Bg
        r_prev = None
        for step in range(max_restarts + 1):
            dX = torch.empty_like(Bg)
            # relative target per column scaled to the original right-hand side
            rn = torch.linalg.norm(rhs, dim=0)
            ratio = float(torch.min(torch.where(rn > 0, bn / torch.where(rn > 0, rn, torch.ones_like(rn)),
                                                torch.full_like(rn, float("inf")))))
            rtol = margin * tol * (ratio if np.isfinite(ratio) else 1.0)
            _lib.check(L.wnfea_pcg_multi(h, system.K.handle, amg.handle, 1, _lib.dptr(system.scale), group,
                                         _lib.dptr(rhs), _lib.dptr(dX), float(rtol), int(max_iter), int(check_every),
                                         its, rel, _lib.stream_of()), "wnfea_pcg_multi")
            iters[g0:g0 + m] += np.array(its[:m])
            Xg += dX
            KX = torch.empty_like(Xg)
            _lib.check(L.wnfea_bsr_spmm(system.K.handle, 0, group, _lib.dptr(Xg), _lib.dptr(KX), _lib.stream_of()),
                       "wnfea_bsr_spmm")
            R = Bg - KX
            r = torch.linalg.norm(R, dim=0) / torch.where(bn > 0, bn, torch.ones_like(bn))
            todo = r > tol
            if r_prev is not None:
                todo &= r <= 0.5 * r_prev  # still improving: not yet at the FP64 floor
            if not bool(todo[:m].any()) or step == max_restarts:
                break
            r_prev = r
            rhs = torch.where(todo[None, :], R, torch.zeros_like(R)).contiguous()
        rels[g0:g0 + m] = r[:m].cpu().numpy()
        U[:, g0:g0 + m] = Xg[:, :m]
    _sync()
    return U, iters, rels, time.perf_counter() - t0


def ir_device(system: SolidSystem, tol: float = 1e-8, inner_rtol: float = 0.05, preconditioner: str = "amg",
              amg=None, amg_precision: str = "fp32", max_outer: int = 30, max_inner: int = 500,
              check_every: int = 16, u=None):
    """FP64 iterative refinement with an FP32 inner PCG on S K S (the
    reference's mixed_fp32 scheme, solid_solver.py:271-300, with an AMG inner
    preconditioner).  Converges when eps_fp32 * kappa_eff of the system is well
    below one.  Returns (u (device FP64), outer steps, total inner iterations,
    ||F-Ku||/||F||, seconds)."""
    import torch
    L = _lib.load()
    dev = system.K.device
    n = system.n_dofs
    pkind = {"none": 0, "jacobi": 1, "amg": 2}[preconditioner]
    dinv = torch.ones(n, dtype=torch.float32, device=dev) if pkind == 1 else None
    zero = u is None
    u = torch.empty(n, dtype=torch.float64, device=dev) if u is None else u
    o = _lib.ctypes.c_int(0)
    i = _lib.ctypes.c_int(0)
    rel = _lib.ctypes.c_double(0.0)
    _sync()
    t0 = time.perf_counter()
    _lib.check(L.wnfea_ir(_Solver.get(), system.K.handle, _lib.dptr(system.scale), pkind,
                          _lib.dptr(dinv) if dinv is not None else None, amg.handle if amg is not None else None,
                          _VT[amg_precision] if pkind == 2 else 1, _lib.dptr(system.F), _lib.dptr(u), float(tol),
                          float(inner_rtol), int(max_outer), int(max_inner), int(check_every), int(zero),
                          _lib.ctypes.byref(o), _lib.ctypes.byref(i), _lib.ctypes.byref(rel), _lib.stream_of()),
               "wnfea_ir")
    _sync()
    return u, o.value, i.value, rel.value, time.perf_counter() - t0


def true_residual(system: SolidSystem, u) -> float:
    """||F - K u|| / ||F|| in FP64 on the device."""
    L = _lib.load()
    out = (_lib.ctypes.c_double * 2)()
    _lib.check(L.wnfea_residual(_Solver.get(), system.K.handle, _lib.dptr(u), _lib.dptr(system.F), None, out,
                                _lib.stream_of()), "wnfea_residual")
    return out[0] / out[1] if out[1] > 0 else 0.0


SOLID_SA_OMEGA = 1.6  # prolongator smoothing factor of the solid hierarchies (see build_amg)


def build_amg(system: SolidSystem, storage: tuple = ("fp64", "fp32", "fp16"), autotune: bool = True, **kw):
    """SA-AMG of the system (rigid-body near-null-space from the node
    coordinates; translations only when the system carries none).  ``storage``
    lists the value precisions the V-cycle may run with.  Unless a launch
    shape is given (waves/unroll/xcd_remap) or ``autotune`` is off, the SpMV
    launch shapes of K's FP64 and FP32 values are autotuned (cached per
    matrix shape) and used by K and by the cycle alike.  The strength
    threshold (``theta`` on the finest level, times ``theta_decay`` per level)
    is mesh-dependent: 0.08 on every level suits the cantilever boxes, the
    1.05M-DOF I-beam converges in 44 instead of 50 iterations with 0.1 and
    decay 0.4 (40 with decay 0.5 and omega 1.6).  The prolongator smoothing P = (I - omega/lambda D^-1 A) P_tent
    takes ``omega`` = 1.6 for solids: on the I-beam 41 iterations against 44
    at the textbook 4/3 (43 at 1.4-1.5 and 1.8, 48 at 1.9), the 64-box batch
    33 against 36, the I-beam's 32 tip load cases 38-48 against 41-51."""
    from .amg_preconditioner import SmoothedAggregationAMG
    kw.setdefault("omega", SOLID_SA_OMEGA)
    tunes = None
    if autotune and not any(k in kw for k in ("waves", "unroll", "xcd_remap")):
        from .autotune import autotune_spmv
        tunes = {p: autotune_spmv(system.K, p) for p in ("fp64", "fp32")}
        kw.update(waves=tunes["fp32"]["waves"], unroll=tunes["fp32"]["unroll"],
                  xcd_remap=tunes["fp32"]["xcd_remap"])
    info = system.K.info()
    need64 = "fp64" in storage and not info["has_fp64_equilibrated"]
    need16 = "fp16" in storage
    if need64 or need16:
        L = _lib.load()
        _lib.check(L.wnfea_bsr_equilibrate(system.K.handle, _lib.dptr(system.K.d_rowidx),
                                           _lib.dptr(system.K.d_colidx), _lib.dptr(system.scale),
                                           int("fp64" in storage), int(need16), _lib.stream_of()),
                   "wnfea_bsr_equilibrate")
    A_hat = system.equilibrated_host()
    dof_scale = (1.0 / system.scale).cpu().numpy()
    amg = SmoothedAggregationAMG(system.K, A_hat, system.coords, system.fixed_mask, dof_scale=dof_scale,
                                 storage=storage, **kw)
    del A_hat
    if tunes:
        t = tunes["fp64"]
        amg.set_tune("fp64", t["waves"], t["unroll"], t["xcd_remap"])
    amg.autotuned = tunes is not None  # attach_matrix_free then retunes the cycle
    return amg


def solve_c3d10_pcg(K, F, rtol: float = 1e-6, max_iter: int = 2000, preconditioner: str = "jacobi",
                    precision: str = "fp64", amg=None) -> tuple[np.ndarray, int, float, float]:
    """Reference-compatible PCG (``solid_solver.py:197-308``) on the device.

    precision "fp64": PCG on K in FP64.  precision "mixed_fp32": the
    reference's refinement loop -- FP64 residual, FP32 inner PCG to
    max(1e-4, rtol) relative with at most max_iter//5 iterations, 10 outer
    steps.  preconditioner: "jacobi", "none", "amg" ("block_jacobi" uses the
    point-Jacobi kernel on the equilibrated system, whose diagonal blocks are
    what block Jacobi inverts up to their off-diagonal entries)."""
    t0 = time.perf_counter()
    system = _as_system(K, F)
    if float(system.F.abs().max().item()) == 0.0:
        return np.zeros(system.n_dofs), 0, 0.0, 0.0
    pc = "jacobi" if preconditioner == "block_jacobi" else preconditioner
    if pc == "amg" and amg is None:
        amg = build_amg(system)
    if precision == "mixed_fp32":
        u, outer, inner, rel, _ = ir_device(system, tol=rtol, inner_rtol=max(1e-4, rtol), preconditioner=pc,
                                            amg=amg, max_outer=10, max_inner=max(max_iter // 5, 1))
        return u.cpu().numpy(), inner, rel, time.perf_counter() - t0
    if pc == "amg":
        x, it, _, _ = _pcg_amg64(system, rtol, max_iter, amg)
    else:
        x, it, _, _ = pcg_device(system, rtol=rtol, max_iter=max_iter, preconditioner=pc, precision="fp64")
    rel = true_residual(system, x)
    return x.cpu().numpy(), it, rel, time.perf_counter() - t0


def _pcg_amg64(system: SolidSystem, rtol, max_iter, amg):
    return pcg_device(system, rtol=rtol, max_iter=max_iter, preconditioner="amg", precision="fp64", amg=amg,
                      amg_precision="fp64")


def solve_c3d10_system(K, F, method: str = "auto", rtol: float = 1e-6, max_iter: int = 2000,
                       preconditioner: str = "jacobi", precision: str = "fp64") -> tuple[np.ndarray, dict]:
    """Reference-compatible entry (``solid_solver.py:132-194``).

    "direct" (and "auto" up to 60 000 DOF, the reference's threshold) has no
    sparse factorisation on the GPU here: it is served by FP64 AMG-PCG,
    restarted on the FP64 true residual until that is 1e-12 or stops falling,
    i.e. down to the round-off level a direct solve reaches."""
    system = _as_system(K, F)
    n = system.n_dofs
    use_direct = method == "direct" or (method == "auto" and n <= 60000)
    if use_direct:
        t0 = time.perf_counter()
        if float(system.F.abs().max().item()) == 0.0:
            return np.zeros(n), {"method": "amg_pcg_fp64_direct_tolerance", "iterations": 0,
                                 "solve_time_sec": 0.0, "rel_res": 0.0}
        amg = build_amg(system, storage=("fp64",))
        try:
            u, it, _, rel, _ = solve_mixed(system, amg, tol=1e-12, amg_precision="fp64",
                                           max_iter=max(max_iter, 5000), margin=0.1, max_restarts=6)
        finally:
            amg.close()
        return u.cpu().numpy(), {"method": "amg_pcg_fp64_direct_tolerance", "iterations": it,
                                 "solve_time_sec": time.perf_counter() - t0, "rel_res": rel}
    u, it, rel, ts = solve_c3d10_pcg(system, None, rtol=rtol, max_iter=max_iter, preconditioner=preconditioner,
                                     precision=precision)
    return u, {"method": f"pcg_{preconditioner}_{precision}", "iterations": it, "solve_time_sec": ts, "rel_res": rel}


def compute_c3d10_stress_field(model: FEAModel, u_3dof, mesh: dict | None = None):
    """(cell_stresses (M,6), cell_vm (M,), nodal_vm (N,)) -- ``solid_solver.py:311-365`` -- on the device."""
    import torch
    L = _lib.load()
    dev = _lib.require_gpu()
    mesh = mesh or mesh_to_device(model, dev)
    M, n = mesh["n_elements"], mesh["n_nodes"]
    U = u_3dof if isinstance(u_3dof, torch.Tensor) else torch.as_tensor(np.asarray(u_3dof, dtype=np.float64))
    U = U.to(device=dev, dtype=torch.float64).contiguous()
    cs = torch.empty((M, 6), dtype=torch.float64, device=dev)
    cv = torch.empty(M, dtype=torch.float64, device=dev)
    nv = torch.empty(n, dtype=torch.float64, device=dev)
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.check(L.wnfea_c3d10_stress(M, n, _lib.dptr(mesh["conn"]), _lib.dptr(mesh["nodes"]), _lib.dptr(mesh["E"]),
                                    _lib.dptr(mesh["nu"]), _lib.dptr(U), _lib.dptr(mesh["n2e_ptr"]),
                                    _lib.dptr(mesh["n2e"]), None, None, _lib.dptr(cs), _lib.dptr(cv), _lib.dptr(nv),
                                    _lib.dptr(bad), _lib.stream_of()), "wnfea_c3d10_stress")
    if int(bad.item()):
        raise ValueError(f"Non-positive Jacobian determinant in C3D10 element {int(bad.item()) - 1}")
    return cs.cpu().numpy(), cv.cpu().numpy(), nv.cpu().numpy()
