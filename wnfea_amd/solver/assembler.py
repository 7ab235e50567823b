"""Beam element matrices and global assembly (drop-in for
``wnfea/solver/assembler.py``).

``assemble_global_system_sparse`` assembles on the device (linear beam
stiffness kernel and C3D10 element kernels, deterministic gathers).  The
single-element helpers ``build_element_stiffness_3d_beam`` (local 12x12
Euler-Bernoulli stiffness, ``assembler.py:13-84``) and
``build_transformation_matrix`` (``assembler.py:87-148``, with the
reference's frame choice for members along a global axis) are closed-form
host utilities for callers that inspect one element; the solve path never
calls them.
"""
from __future__ import annotations

import numpy as np

from .beam_solver import assemble_global_system_sparse  # noqa: F401  (device assembly)


def build_element_stiffness_3d_beam(node1, node2, E: float, G: float, A: float, Iy: float, Iz: float,
                                    J: float) -> np.ndarray:
    """Local 12x12 stiffness, DOFs [u, v, w, rx, ry, rz] per node."""
    L = float(np.linalg.norm(np.asarray(node2, float) - np.asarray(node1, float)))
    K = np.zeros((12, 12))
    if L < 1e-12:
        return K

    def pair(i, j, a, b):  # K[i,i] = K[j,j] = a, K[i,j] = K[j,i] = b
        K[i, i] = K[j, j] = a
        K[i, j] = K[j, i] = b

    pair(0, 6, E * A / L, -E * A / L)         # axial
    pair(3, 9, G * J / L, -G * J / L)         # torsion
    for (v, r, I, s) in ((1, 5, Iz, 1.0), (2, 4, Iy, -1.0)):  # bending in the xy and xz planes
        k1, k2, k3, k4 = 12 * E * I / L ** 3, 6 * E * I / L ** 2, 4 * E * I / L, 2 * E * I / L
        vv, rr = (v, v + 6), (r, r + 6)
        pair(vv[0], vv[1], k1, -k1)
        pair(rr[0], rr[1], k3, k4)
        for a, sgn in ((rr[0], 1.0), (rr[1], 1.0)):
            K[vv[0], a] = K[a, vv[0]] = s * sgn * k2
            K[vv[1], a] = K[a, vv[1]] = -s * sgn * k2
    return K


def _frame(d: np.ndarray) -> np.ndarray:
    """Rows lx, ly, lz of the element frame (the reference's axis rules)."""
    L = float(np.linalg.norm(d))
    for ax, (ly, lz) in ((0, ((0, 1, 0), (0, 0, 1))), (1, ((1, 0, 0), (0, 0, 1))), (2, ((1, 0, 0), (0, 1, 0)))):
        if np.isclose(abs(d[ax]), L):
            lx = np.zeros(3)
            lx[ax] = 1.0 if d[ax] > 0 else -1.0
            return np.array([lx, ly, lz], dtype=float)
    lx = d / L
    ly = np.cross((0.0, 0.0, 1.0), lx)
    if np.linalg.norm(ly) < 1e-9:
        ly = np.cross((0.0, 1.0, 0.0), lx)
    ly /= np.linalg.norm(ly)
    lz = np.cross(lx, ly)
    return np.array([lx, ly, lz / np.linalg.norm(lz)])


def build_transformation_matrix(node1, node2) -> np.ndarray:
    """12x12 block-diagonal rotation with K_global = T^T K_local T."""
    d = np.asarray(node2, float) - np.asarray(node1, float)
    if np.linalg.norm(d) < 1e-12:
        return np.eye(12)
    return np.kron(np.eye(4), _frame(d))
