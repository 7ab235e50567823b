"""Equilibrium residual of beam, solid and mixed models (drop-in for
``wnfea/solver/residual.py:20-166``).

The internal forces come from the device kernels (co-rotational beams and
matrix-free C3D10 solids, 6 DOF per node, deterministic gathers); the device
copies of a model's mesh are kept between calls (keyed by the model and its
mesh arrays), so a Krylov loop that evaluates the residual once per
iteration pays the upload once.  Load and support bookkeeping is host code,
as in the reference.
"""
from __future__ import annotations

import weakref

import numpy as np

from .. import _lib
from ..boundary.conditions import DOFType
from .dof_manager import DOFManager


def _mesh_node(model, item, n_nodes: int):
    """Mesh node of a load or support: geometry-node ids go through the
    model's geometry-to-mesh map when it has one."""
    nid = item.node_id
    if getattr(item, "is_geometry_node", False) and getattr(model, "geometry_to_mesh_node_map", None):
        nid = model.geometry_to_mesh_node_map.get(nid)
    if nid is None or nid >= n_nodes:
        return None
    return int(nid)


def build_external_force_vector(model, dof_mgr: DOFManager | None = None) -> np.ndarray:
    """F_ext over 6 DOF per node, or condensed to the active DOFs with a
    DOF manager (``residual.py:20-47``)."""
    n = len(model.mesh_nodes)
    F = np.zeros(6 * n)
    for ld in model.loads:
        nid = _mesh_node(model, ld, n)
        if nid is not None:
            F[6 * nid:6 * nid + 6] += np.asarray(ld.force_vector, dtype=np.float64)
    return dof_mgr.condense_forces(F) if dof_mgr is not None else F


def get_boundary_constraints(model, dof_mgr: DOFManager | None = None):
    """(constrained DOFs, prescribed values), active numbering with a DOF
    manager (``residual.py:50-95``)."""
    n = len(model.mesh_nodes)
    dofs, vals = [], []
    for s in model.supports:
        nid = _mesh_node(model, s, n)
        if nid is None:
            continue
        for d, c in enumerate(s.constraints):
            g = dof_mgr.node_dof_to_global.get((nid, d)) if dof_mgr is not None else 6 * nid + d
            if g is None:
                continue
            if c.dof_type == DOFType.FIXED:
                dofs.append(g)
                vals.append(0.0)
            elif c.dof_type == DOFType.PRESCRIBED:
                dofs.append(g)
                vals.append(float(c.value))
    return np.array(dofs, dtype=np.int64), np.array(vals, dtype=np.float64)


class _ForceContext:
    """Device copies of a model's beams and solids for the force kernels."""

    def __init__(self, model):
        import torch
        from ..mesh.lattice import beam_props
        from .solid_solver import mesh_to_device
        L = _lib.load()
        dev = _lib.require_gpu()
        self.key = self._key(model)
        nodes = np.ascontiguousarray(model.mesh_nodes, dtype=np.float64)
        self.n = len(nodes)
        self.beams = None
        if model.mesh_elements is not None and len(model.mesh_elements):
            conn = np.ascontiguousarray(model.mesh_elements, dtype=np.int32)
            M = len(conn)
            ptr = np.zeros(self.n + 1, dtype=np.int32)
            inc = np.zeros(2 * M, dtype=np.int32)
            _lib.check(L.wnfea_node_incidence(self.n, M, 2, _lib.hptr(conn), _lib.hptr(ptr), _lib.hptr(inc)),
                       "wnfea_node_incidence")
            t = lambda a: torch.as_tensor(a, device=dev)  # noqa: E731
            props = np.ascontiguousarray(beam_props(model))
            self.beams = {"M": M, "conn": t(conn), "nodes": t(nodes), "props": t(props), "ptr": t(ptr),
                          "inc": t(inc), "fe": torch.empty((M, 12), dtype=torch.float64, device=dev)}
        self.solids = None
        if getattr(model, "solid_elements", None) is not None and len(model.solid_elements):
            m = mesh_to_device(model, dev)
            m["fe"] = torch.empty((m["n_elements"], 30), dtype=torch.float64, device=dev)
            self.solids = m

    @staticmethod
    def _key(model):
        arrs = (model.mesh_nodes, model.mesh_elements, getattr(model, "solid_elements", None))
        return tuple((id(a), None if a is None else len(a)) for a in arrs)

    def forces(self, U):
        """F_int (device, 6 per node) of device U (6 per node)."""
        import torch
        L = _lib.load()
        F = torch.zeros(6 * self.n, dtype=torch.float64, device=U.device)
        acc = 0
        if self.beams is not None:
            b = self.beams
            _lib.check(L.wnfea_beam_forces(b["M"], self.n, _lib.dptr(b["conn"]), _lib.dptr(b["nodes"]),
                                           _lib.dptr(b["props"]), _lib.dptr(U), _lib.dptr(b["ptr"]),
                                           _lib.dptr(b["inc"]), _lib.dptr(b["fe"]), _lib.dptr(F), 0,
                                           _lib.stream_of()), "wnfea_beam_forces")
            acc = 1
        if self.solids is not None:
            m = self.solids
            _lib.check(L.wnfea_c3d10_fint(m["n_elements"], self.n, _lib.dptr(m["conn"]), _lib.dptr(m["nodes"]),
                                          _lib.dptr(m["E"]), _lib.dptr(m["nu"]), _lib.dptr(U), 6,
                                          _lib.dptr(m["n2e_ptr"]), _lib.dptr(m["n2e"]), _lib.dptr(m["fe"]),
                                          _lib.dptr(F), acc, _lib.stream_of()), "wnfea_c3d10_fint")
        return F


_CTX: dict = {}  # id(model) -> (weak reference to the model, context)


def _context(model) -> _ForceContext:
    hit = _CTX.get(id(model))
    if hit is not None and hit[0]() is model and hit[1].key == _ForceContext._key(model):
        return hit[1]
    ctx = _ForceContext(model)
    _CTX[id(model)] = (weakref.ref(model, lambda _r, k=id(model): _CTX.pop(k, None)), ctx)
    return ctx


def compute_internal_forces(model, U_full, device: str = "hip", block_size: int = 0, grid_size: int = 0):
    """F_int(U) over 6 DOF per node for beams, C3D10 solids (translations of
    their nodes) or both (``residual.py:98-122``).  numpy in -> numpy out,
    device tensor in -> device tensor out; ``device`` and the launch-shape
    arguments are accepted for API parity (the kernels always run on the GPU)."""
    import torch
    on_dev = isinstance(U_full, torch.Tensor)
    U = (U_full if on_dev else torch.as_tensor(np.asarray(U_full, dtype=np.float64), device="cuda"))
    F = _context(model).forces(U.to(torch.float64).contiguous())
    return F if on_dev else F.cpu().numpy()


def compute_equilibrium_residual(model, U, F_ext, load_factor: float = 1.0, constrained_dofs=None,
                                 prescribed_vals=None, dof_mgr: DOFManager | None = None, device: str = "hip"):
    """R = F_int(U) - load_factor F_ext in the active DOFs, R_c = U_c -
    load_factor * prescribed on constrained DOFs (``residual.py:125-166``)."""
    if dof_mgr is None:
        dof_mgr = DOFManager(model)
    if constrained_dofs is None or prescribed_vals is None:
        constrained_dofs, prescribed_vals = get_boundary_constraints(model, dof_mgr)
    U = np.asarray(U, dtype=np.float64)
    U_full = dof_mgr.expand_displacements(U) if len(U) == dof_mgr.total_active_dofs else U
    R = dof_mgr.condense_forces(compute_internal_forces(model, U_full)) - load_factor * np.asarray(F_ext)
    if len(constrained_dofs):
        R[constrained_dofs] = U[constrained_dofs] - load_factor * prescribed_vals
    return R
