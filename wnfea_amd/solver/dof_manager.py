"""Active-DOF bookkeeping for beam (6 DOF/node), solid (3 DOF/node) and mixed
models with rigid couplings (drop-in for ``wnfea/solver/dof_manager.py:17-171``).

Same numbering as the reference: nodes in order; a beam node (or any node of
a model without solids) has 6 active DOFs, a pure solid node its 3
translations, a coupling slave none (its motion follows the master,
u_s = T_s [u_m; theta_m]).  The maps are built with numpy index arithmetic,
so a 400 k-node model is set up in milliseconds; ``projection`` gives the
sparse u_full = T u_active map (``build_projection_matrix`` the dense one of
the reference).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from ..boundary.coupling import coupling_matrices


class DOFManager:
    def __init__(self, model):
        self.model = model
        nodes = np.asarray(model.mesh_nodes, dtype=np.float64)
        n = self.n_nodes = len(nodes)
        beams = model.mesh_elements if model.mesh_elements is not None else np.zeros((0, 2), np.int64)
        solids = (model.solid_elements if getattr(model, "solid_elements", None) is not None
                  else np.zeros((0, 10), np.int64))
        is_beam = np.zeros(n, dtype=bool)
        is_solid = np.zeros(n, dtype=bool)
        if len(beams):
            is_beam[np.asarray(beams, dtype=np.int64).ravel()] = True
        if len(solids):
            is_solid[np.asarray(solids, dtype=np.int64).ravel()] = True
        self.beam_nodes = set(np.flatnonzero(is_beam).tolist())
        self.solid_nodes = set(np.flatnonzero(is_solid).tolist())
        # couplings: slave -> (master, T)
        self.slave_to_coupling: dict = {}
        is_slave = np.zeros(n, dtype=bool)
        self._coup = []  # (master, slaves (k,), T (k,3,6))
        for c in getattr(model, "couplings", []) or []:
            sl = np.asarray(c.slave_node_ids, dtype=np.int64)
            T = coupling_matrices(nodes[c.master_node_id], nodes[sl])
            self._coup.append((int(c.master_node_id), sl, T))
            is_slave[sl] = True
            for s, Ti in zip(sl.tolist(), T):
                self.slave_to_coupling[s] = (int(c.master_node_id), Ti)
        self.slave_nodes = set(np.flatnonzero(is_slave).tolist())
        # active DOFs per node: 6 for beam nodes (and every node of a solid-free
        # model), 3 for pure solid nodes, none for slaves
        six = is_beam | (~is_solid & (len(self.solid_nodes) == 0))
        count = np.where(is_slave, 0, np.where(six, 6, 3))
        start = np.concatenate([[0], np.cumsum(count)])
        self.total_active_dofs = int(start[-1])
        node_of = np.repeat(np.arange(n), count)
        local = np.arange(self.total_active_dofs) - start[node_of]
        self._full_index = node_of * 6 + local          # active DOF -> full (6 per node) index
        self._active_of_full = np.full(6 * n, -1, dtype=np.int64)
        self._active_of_full[self._full_index] = np.arange(self.total_active_dofs)
        self.global_to_node_dof = list(zip(node_of.tolist(), local.tolist()))
        self.node_dof_to_global = {nd: g for g, nd in enumerate(self.global_to_node_dof)}
        self.active_dofs_per_node = {i: list(range(int(c))) for i, c in enumerate(count)}

    # ---------------------------------------------------------------- maps
    def projection(self) -> sp.csr_matrix:
        """Sparse T (6 n_nodes x total_active_dofs) with u_full = T u_active."""
        rows = [self._full_index]
        cols = [np.arange(self.total_active_dofs)]
        vals = [np.ones(self.total_active_dofs)]
        for m, sl, T in self._coup:
            mg = self._active_of_full[6 * m + np.arange(6)]
            ok = mg >= 0
            k = len(sl)
            r = (6 * sl[:, None, None] + np.arange(3)[None, :, None]) * np.ones((1, 1, 6), dtype=np.int64)
            c = np.broadcast_to(mg[None, None, :], (k, 3, 6))
            rows.append(r[:, :, ok].ravel())
            cols.append(c[:, :, ok].ravel())
            vals.append(T[:, :, ok].ravel())
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(6 * self.n_nodes, self.total_active_dofs))

    def build_projection_matrix(self) -> np.ndarray:
        """Dense T (``dof_manager.py:141-171``)."""
        return self.projection().toarray()

    def expand_displacements(self, u_active) -> np.ndarray:
        """u_full (6 n_nodes) from u_active; slaves follow their masters, with
        zero rotations (``dof_manager.py:82-108``)."""
        u = np.zeros(6 * self.n_nodes)
        u[self._full_index] = np.asarray(u_active, dtype=np.float64)
        for m, sl, T in self._coup:
            um = u[6 * m:6 * m + 6]
            us = T @ um  # (k, 3)
            U = u.reshape(-1, 6)
            U[sl, :3] = us
            U[sl, 3:] = 0.0
        return u

    def condense_forces(self, f_full) -> np.ndarray:
        """Active-DOF forces: slave forces and their moments about the master
        added to the master, then the active entries (``dof_manager.py:110-139``)."""
        f = np.asarray(f_full, dtype=np.float64).reshape(-1, 6).copy()
        nodes = np.asarray(self.model.mesh_nodes, dtype=np.float64)
        for m, sl, _ in self._coup:
            fs = f[sl, :3]
            r = nodes[sl] - nodes[m]
            f[m, :3] += fs.sum(axis=0)
            f[m, 3:] += np.cross(r, fs).sum(axis=0)
        return f.reshape(-1)[self._full_index]
