"""Co-rotational 3D beam element (drop-in for
``wnfea/solver/corotational_beam.py:15-233``).

``compute_corotational_element_forces`` runs the device force kernel on the
one element (the kernel every solve uses for all elements at once).
``compute_element_triad`` (the current co-rotating frame, chord length and
initial length, ``:15-123``) and ``build_corotational_projector`` (the 6x12
map from global element DOFs to the local deformation modes, ``:126-170``)
are closed-form host utilities for callers that inspect one element.
"""
from __future__ import annotations

import numpy as np

from .. import _lib

_AXIS_FRAMES = {0: ((0.0, 1.0, 0.0), (0.0, 0.0, 1.0)), 1: ((1.0, 0.0, 0.0), (0.0, 0.0, 1.0)),
                2: ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0))}


def _rotate(v, axis, c, s):
    """Rodrigues rotation of v about the unit axis by the angle with cosine c, sine s."""
    return v * c + np.cross(axis, v) * s + axis * float(axis @ v) * (1.0 - c)


def compute_element_triad(node1, node2, u_e=None):
    """(r1, r2, r3, L, L0): the rigidly rotated initial triad (the chord
    rotation, then the mean nodal twist about the chord), re-orthonormalised."""
    X1, X2 = np.asarray(node1, float), np.asarray(node2, float)
    d0 = X2 - X1
    L0 = float(np.linalg.norm(d0))
    if L0 < 1e-12:
        raise ValueError("Element has zero undeformed length")
    e1 = d0 / L0
    for ax, (a2, a3) in _AXIS_FRAMES.items():
        if np.isclose(abs(e1[ax]), 1.0):
            e2, e3 = np.array(a2), np.array(a3)
            break
    else:
        e2 = np.cross((0.0, 0.0, 1.0), e1)
        if np.linalg.norm(e2) < 1e-9:
            e2 = np.cross((0.0, 1.0, 0.0), e1)
        e2 = e2 / np.linalg.norm(e2)
        e3 = np.cross(e1, e2)
        e3 = e3 / np.linalg.norm(e3)
    if u_e is None or np.linalg.norm(u_e) < 1e-15:
        return e1, e2, e3, L0, L0
    u = np.asarray(u_e, float)
    d = d0 + u[6:9] - u[0:3]
    L = float(np.linalg.norm(d))
    if L < 1e-12:
        raise ValueError("Element collapsed to zero length during deformation")
    r1 = d / L
    ax = np.cross(e1, r1)
    s = float(np.linalg.norm(ax))
    c = float(e1 @ r1)
    if s > 1e-12:
        ax = ax / s
        phi = np.arctan2(s, c)
        r2, r3 = _rotate(e2, ax, np.cos(phi), np.sin(phi)), _rotate(e3, ax, np.cos(phi), np.sin(phi))
    elif c < 0.0:
        r2, r3 = -e2, -e3
    else:
        r2, r3 = e2.copy(), e3.copy()
    tw = 0.5 * float((u[3:6] + u[9:12]) @ r1)
    if abs(tw) > 1e-12:
        ct, st = np.cos(tw), np.sin(tw)
        r2, r3 = r2 * ct + np.cross(r1, r2) * st, r3 * ct + np.cross(r1, r3) * st
    r1 = r1 / np.linalg.norm(r1)
    r2 = r2 - (r2 @ r1) * r1
    r2 = r2 / np.linalg.norm(r2)
    r3 = np.cross(r1, r2)
    return r1, r2, r3 / np.linalg.norm(r3), L, L0


def build_corotational_projector(r1, r2, r3, L: float) -> np.ndarray:
    """6x12 B with delta p = B delta u_e, p = [u_l, th_y1, th_z1, th_x, th_y2, th_z2]."""
    r1, r2, r3 = (np.asarray(v, float) for v in (r1, r2, r3))
    B = np.zeros((6, 12))
    t1, t2 = slice(0, 3), slice(6, 9)   # translations of node 1, node 2
    q1, q2 = slice(3, 6), slice(9, 12)  # rotations
    B[0, t1], B[0, t2] = -r1, r1
    B[3, q1], B[3, q2] = -r1, r1
    for row, q in ((1, q1), (4, q2)):   # bending y at node 1 / node 2
        B[row, t1], B[row, t2], B[row, q] = -r3 / L, r3 / L, r2
    for row, q in ((2, q1), (5, q2)):   # bending z
        B[row, t1], B[row, t2], B[row, q] = r2 / L, -r2 / L, r3
    return B


def compute_corotational_element_forces(node1, node2, u_e, E: float, G: float, A: float, Iy: float, Iz: float,
                                        J: float) -> np.ndarray:
    """12-DOF internal force vector of one element, computed on the device."""
    import torch
    L = _lib.load()
    dev = _lib.require_gpu()
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)  # noqa: E731
    nodes = t(np.array([node1, node2], dtype=float))
    conn = t(np.array([[0, 1]]), torch.int32)
    props = t(np.array([[E, G, A, Iy, Iz, J]], dtype=float))
    U = t(np.asarray(u_e, float).reshape(12))
    ptr = t(np.array([0, 1, 2]), torch.int32)
    inc = t(np.array([0, 1]), torch.int32)
    fe = torch.empty((1, 12), dtype=torch.float64, device=dev)
    F = torch.empty(12, dtype=torch.float64, device=dev)
    _lib.check(L.wnfea_beam_forces(1, 2, _lib.dptr(conn), _lib.dptr(nodes), _lib.dptr(props), _lib.dptr(U),
                                   _lib.dptr(ptr), _lib.dptr(inc), _lib.dptr(fe), _lib.dptr(F), 0, _lib.stream_of()),
               "wnfea_beam_forces")
    return F.cpu().numpy()
