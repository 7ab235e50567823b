"""Reverse Cuthill-McKee renumbering (drop-in for ``wnfea/mesh/rcm.py``).

Same functions, arguments and results as the reference (``rcm.py:20-196``).
The node graph and the RCM breadth-first search run in the native library
(``wnfea_node_graph``, ``wnfea_rcm``); the seed order is numpy's ``argsort``
of the int32 node degrees, exactly the call scipy's reverse_cuthill_mckee
makes inside the reference (``rcm.py:85``), so the permutation is the one the
reference produces on the same host.
"""
from __future__ import annotations

import copy

import numpy as np
import scipy.sparse as sp

from .. import _lib
from ..model import FEAModel


def _graph(n_nodes: int, conn: np.ndarray, npe: int):
    L = _lib.load()
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    if conn.size and (conn.min() < 0 or conn.max() >= n_nodes):
        raise ValueError("element connectivity references a node outside the mesh")
    rowptr = np.zeros(n_nodes + 1, dtype=np.int32)
    m = len(conn)
    cap = min(n_nodes * n_nodes, m * npe * npe + n_nodes)
    colidx = np.zeros(max(cap, 1), dtype=np.int32)
    nnz = L.wnfea_node_graph(n_nodes, m, npe, _lib.hptr(conn), None, _lib.hptr(rowptr), _lib.hptr(colidx), cap)
    if nnz < 0:
        raise RuntimeError(f"wnfea_node_graph failed ({nnz})")
    return rowptr, colidx[:nnz].copy()


def compute_node_adjacency_matrix(
    n_nodes: int,
    beam_elements: np.ndarray | None = None,
    solid_elements: np.ndarray | None = None,
) -> sp.csr_matrix:
    """Symmetric int8 CSR adjacency with self-loops (``rcm.py:20-70``)."""
    parts = []
    if beam_elements is not None and len(beam_elements) > 0:
        parts.append(_graph(n_nodes, np.asarray(beam_elements).reshape(-1, 2), 2))
    if solid_elements is not None and len(solid_elements) > 0:
        parts.append(_graph(n_nodes, np.asarray(solid_elements).reshape(-1, 10), 10))
    if not parts:
        return sp.identity(n_nodes, dtype=np.int8, format="csr")
    mats = [sp.csr_matrix((np.ones(len(c), dtype=np.int8), c, r), shape=(n_nodes, n_nodes)) for r, c in parts]
    adj = mats[0]
    for m in mats[1:]:
        adj = (adj + m).tocsr()
    adj.sum_duplicates()
    adj.data[:] = 1
    return adj


def rcm_seed_order(adj: sp.csr_matrix) -> np.ndarray:
    """numpy argsort of the scipy-style degrees (entries + 1 for a stored diagonal)."""
    n = adj.shape[0]
    deg = np.diff(adj.indptr).astype(np.int32)
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(adj.indptr))
    has_diag = np.zeros(n, dtype=np.int32)
    has_diag[rows[rows == adj.indices]] = 1
    return np.argsort(deg + has_diag).astype(np.int32)


def compute_rcm_order(adj: sp.csr_matrix) -> tuple[np.ndarray, np.ndarray]:
    """(new_to_old, old_to_new) of the reverse Cuthill-McKee order (``rcm.py:73-92``)."""
    A = sp.csr_matrix(adj)
    A.sort_indices()
    n = A.shape[0]
    ptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
    ind = np.ascontiguousarray(A.indices, dtype=np.int32)
    seed = np.ascontiguousarray(rcm_seed_order(A))
    new_to_old = np.zeros(n, dtype=np.int32)
    _lib.check(_lib.load().wnfea_rcm(n, _lib.hptr(ptr), _lib.hptr(ind), _lib.hptr(seed), _lib.hptr(new_to_old)),
               "wnfea_rcm")
    old_to_new = np.empty(n, dtype=np.int32)
    old_to_new[new_to_old] = np.arange(n, dtype=np.int32)
    return new_to_old, old_to_new


def compute_matrix_bandwidth(A: sp.spmatrix) -> int:
    """Half-bandwidth max |i - j| over the non-zeros (``rcm.py:95-102``)."""
    r, c = A.nonzero()
    if len(r) == 0:
        return 0
    return int(np.max(np.abs(r - c)))


def apply_rcm_to_model(model: FEAModel) -> tuple[FEAModel, np.ndarray, np.ndarray]:
    """Permuted shallow copy of the model plus (new_to_old, old_to_new) (``rcm.py:105-163``)."""
    if model.mesh_nodes is None or len(model.mesh_nodes) == 0:
        raise ValueError("Cannot apply RCM to an unmeshed model.")
    n = len(model.mesh_nodes)
    adj = compute_node_adjacency_matrix(n, model.mesh_elements, getattr(model, "solid_elements", None))
    new_to_old, old_to_new = compute_rcm_order(adj)
    out = copy.copy(model)
    out.mesh_nodes = model.mesh_nodes[new_to_old].copy()
    if model.mesh_elements is not None and len(model.mesh_elements) > 0:
        out.mesh_elements = old_to_new[model.mesh_elements].copy()
    if getattr(model, "solid_elements", None) is not None and len(model.solid_elements) > 0:
        out.solid_elements = old_to_new[model.solid_elements].copy()
    out.supports = []
    for s in model.supports:
        c = copy.deepcopy(s)
        if not c.is_geometry_node:
            c.node_id = int(old_to_new[c.node_id])
        out.supports.append(c)
    out.loads = []
    for ld in model.loads:
        c = copy.deepcopy(ld)
        if not c.is_geometry_node:
            c.node_id = int(old_to_new[c.node_id])
        out.loads.append(c)
    return out, new_to_old, old_to_new


def permute_solution_to_original(u_permuted: np.ndarray, new_to_old: np.ndarray,
                                 dofs_per_node: int = 6) -> np.ndarray:
    """Scatter an RCM-ordered solution back to the user numbering (``rcm.py:166-196``)."""
    u = np.asarray(u_permuted).ravel()
    n = len(new_to_old)
    if len(u) != n * dofs_per_node:
        raise ValueError(f"Length of u_permuted ({len(u)}) does not match {n} nodes * {dofs_per_node} DOFs.")
    idx = (np.asarray(new_to_old, dtype=np.int64)[:, None] * dofs_per_node + np.arange(dofs_per_node)).ravel()
    out = np.empty_like(u)
    out[idx] = u
    return out
