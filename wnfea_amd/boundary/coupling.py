"""Rigid coupling of 3-DOF slave nodes to a 6-DOF master node (drop-in for
``wnfea/boundary/coupling.py:16-69``).

Slave i moves with the master's translation and rotation:
u_i = u_m + theta_m x r_i with r_i = X_i - X_m, i.e. u_i = T_i [u_m; theta_m]
with T_i = [I, -[r_i]x]; the slave forces go to the master as F_m = sum f_i,
M_m = sum r_i x f_i (the transpose of the same map)."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def coupling_matrices(master_coord, slave_coords) -> np.ndarray:
    """(k, 3, 6) stack of T_i = [I, -[r_i]x] for the k slaves."""
    r = np.asarray(slave_coords, dtype=np.float64).reshape(-1, 3) - np.asarray(master_coord, dtype=np.float64)
    T = np.zeros((len(r), 3, 6))
    T[:, 0, 0] = T[:, 1, 1] = T[:, 2, 2] = 1.0
    # -[r]x: row-wise  (0, rz, -ry), (-rz, 0, rx), (ry, -rx, 0)
    T[:, 0, 4], T[:, 0, 5] = r[:, 2], -r[:, 1]
    T[:, 1, 3], T[:, 1, 5] = -r[:, 2], r[:, 0]
    T[:, 2, 3], T[:, 2, 4] = r[:, 1], -r[:, 0]
    return T


@dataclass
class RigidCoupling:
    master_node_id: int
    slave_node_ids: list = field(default_factory=list)
    label: str = ""

    def compute_coupling_matrices(self, master_coord, slave_coords) -> list:
        """List of the 3x6 T_i (``coupling.py:33-69``)."""
        return list(coupling_matrices(master_coord, slave_coords))
