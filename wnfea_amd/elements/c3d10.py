"""C3D10 element API (drop-in for ``wnfea/elements/c3d10.py``).

``element_stiffness_c3d10``, ``element_internal_forces_c3d10`` and
``element_stresses_c3d10`` run the gfx950 element kernels (one element or a
batch); their results match the reference functions of the same names
(``c3d10.py:219-318``).  The per-point helpers ``shape_functions_c3d10``,
``jacobian_c3d10``, ``b_matrix_c3d10`` and ``elasticity_matrix_3d`` are the
reference's closed-form formulas (``c3d10.py:35-216``) kept as host utilities
for callers that inspect a single point; the solve path never calls them.
"""
from __future__ import annotations

import numpy as np

from .. import _lib

_ALPHA = (5.0 - np.sqrt(5.0)) / 20.0
_BETA = (5.0 + 3.0 * np.sqrt(5.0)) / 20.0
_WEIGHT = 1.0 / 24.0
GAUSS_POINTS_C3D10 = (
    np.array([_ALPHA, _ALPHA, _ALPHA]),
    np.array([_BETA, _ALPHA, _ALPHA]),
    np.array([_ALPHA, _BETA, _ALPHA]),
    np.array([_ALPHA, _ALPHA, _BETA]),
)
GAUSS_WEIGHTS_C3D10 = (_WEIGHT, _WEIGHT, _WEIGHT, _WEIGHT)


def shape_functions_c3d10(xi: float, eta: float, zeta: float) -> tuple[np.ndarray, np.ndarray]:
    """N (10,) and dN/d(xi,eta,zeta) (3,10) in Abaqus node order."""
    L = (1.0 - xi - eta - zeta, xi, eta, zeta)
    N = np.array([L[0] * (2 * L[0] - 1), L[1] * (2 * L[1] - 1), L[2] * (2 * L[2] - 1), L[3] * (2 * L[3] - 1),
                  4 * L[0] * L[1], 4 * L[1] * L[2], 4 * L[2] * L[0], 4 * L[0] * L[3], 4 * L[1] * L[3],
                  4 * L[2] * L[3]])
    # derivative of each N_i with respect to each volume coordinate L_k
    dL = np.zeros((4, 10))
    for k in range(4):
        dL[k, k] = 4 * L[k] - 1
    for m, (a, b) in enumerate(((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))):
        dL[a, 4 + m] = 4 * L[b]
        dL[b, 4 + m] = 4 * L[a]
    dN = dL[1:4] - dL[0]
    return N, dN


def jacobian_c3d10(node_coords: np.ndarray, dN_dxi: np.ndarray) -> tuple[np.ndarray, float, np.ndarray]:
    J = dN_dxi @ np.asarray(node_coords, dtype=np.float64)
    det = float(np.linalg.det(J))
    if det <= 0.0:
        raise ValueError(f"Non-positive Jacobian determinant in C3D10 element (det_J = {det:.4e}). "
                         "Element is inverted, degenerate, or has bad node numbering.")
    return J, det, np.linalg.inv(J)


def b_matrix_c3d10(node_coords: np.ndarray, xi: float, eta: float, zeta: float) -> tuple[np.ndarray, float]:
    _, dN = shape_functions_c3d10(xi, eta, zeta)
    _, det, inv = jacobian_c3d10(node_coords, dN)
    g = inv @ dN
    B = np.zeros((6, 30))
    c = 3 * np.arange(10)
    B[0, c], B[1, c + 1], B[2, c + 2] = g[0], g[1], g[2]
    B[3, c], B[3, c + 1] = g[1], g[0]
    B[4, c + 1], B[4, c + 2] = g[2], g[1]
    B[5, c], B[5, c + 2] = g[2], g[0]
    return B, det


def elasticity_matrix_3d(E: float, nu: float) -> np.ndarray:
    f = E / ((1.0 + nu) * (1.0 - 2.0 * nu))
    C = np.full((3, 3), f * nu)
    np.fill_diagonal(C, f * (1.0 - nu))
    out = np.zeros((6, 6))
    out[:3, :3] = C
    out[3:, 3:] = np.eye(3) * (E / (2.0 * (1.0 + nu)))
    return out


# ---------------------------------------------------------------- GPU element kernels
def _batch(coords, E, nu):
    import torch
    dev = _lib.require_gpu()
    X = np.asarray(coords, dtype=np.float64)
    single = X.ndim == 2
    X = X.reshape(-1, 10, 3)
    M = X.shape[0]
    nodes = torch.as_tensor(np.ascontiguousarray(X.reshape(-1, 3)), device=dev)
    conn = torch.arange(M * 10, dtype=torch.int32, device=dev).reshape(M, 10)
    Et = torch.as_tensor(np.broadcast_to(np.asarray(E, dtype=np.float64), (M,)).copy(), device=dev)
    nut = torch.as_tensor(np.broadcast_to(np.asarray(nu, dtype=np.float64), (M,)).copy(), device=dev)
    return single, M, nodes, conn, Et, nut, dev


def _raise_bad(bad: int, what: str):
    if bad:
        raise ValueError(f"Non-positive Jacobian determinant in C3D10 element {bad - 1} ({what}). "
                         "Element is inverted, degenerate, or has bad node numbering.")


def element_stiffness_c3d10(node_coords: np.ndarray, E: float, nu: float) -> np.ndarray:
    """Ke (30,30) for coords (10,3), or (M,30,30) for coords (M,10,3); GPU FP64."""
    import torch
    X = np.asarray(node_coords, dtype=np.float64)
    if X.shape[-2:] != (10, 3):
        raise ValueError(f"Expected node_coords of shape (10, 3), got {X.shape}")
    single, M, nodes, conn, Et, nut, dev = _batch(X, E, nu)
    Ke = torch.empty((M, 30, 30), dtype=torch.float64, device=dev)
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    L = _lib.load()
    _lib.check(L.wnfea_c3d10_ke(M, _lib.dptr(conn), _lib.dptr(nodes), _lib.dptr(Et), _lib.dptr(nut),
                                _lib.dptr(Ke), _lib.dptr(bad), _lib.stream_of()), "wnfea_c3d10_ke")
    _raise_bad(int(bad.item()), "stiffness")
    out = Ke.cpu().numpy()
    return out[0] if single else out


def element_internal_forces_c3d10(node_coords: np.ndarray, u_e: np.ndarray, E: float, nu: float) -> np.ndarray:
    """f_int (30,) = Ke u_e matrix-free (or (M,30) for a batch); GPU FP64."""
    import torch
    single, M, nodes, conn, Et, nut, dev = _batch(node_coords, E, nu)
    U = torch.as_tensor(np.ascontiguousarray(np.asarray(u_e, dtype=np.float64).reshape(M * 10, 3)), device=dev)
    # element-local numbering: each element owns its 10 nodes, so the
    # node-wise gather is the identity incidence
    n = M * 10
    ptr = torch.arange(n + 1, dtype=torch.int32, device=dev)
    inc = torch.arange(n, dtype=torch.int32, device=dev)
    fe = torch.empty((M, 30), dtype=torch.float64, device=dev)
    F = torch.zeros(n * 3, dtype=torch.float64, device=dev)
    L = _lib.load()
    _lib.check(L.wnfea_c3d10_fint(M, n, _lib.dptr(conn), _lib.dptr(nodes), _lib.dptr(Et), _lib.dptr(nut),
                                  _lib.dptr(U), 3, _lib.dptr(ptr), _lib.dptr(inc), _lib.dptr(fe), _lib.dptr(F), 0,
                                  _lib.stream_of()), "wnfea_c3d10_fint")
    out = fe.cpu().numpy()
    return out[0] if single else out


def element_stresses_c3d10(node_coords: np.ndarray, u_e: np.ndarray, E: float, nu: float):
    """(gp_stresses (4,6), gp_von_mises (4,), max von Mises) as ``c3d10.py:286-318``; GPU FP64."""
    import torch
    single, M, nodes, conn, Et, nut, dev = _batch(node_coords, E, nu)
    U = torch.as_tensor(np.ascontiguousarray(np.asarray(u_e, dtype=np.float64).reshape(M * 10, 3)), device=dev)
    gs = torch.empty((M, 4, 6), dtype=torch.float64, device=dev)
    gv = torch.empty((M, 4), dtype=torch.float64, device=dev)
    cs = torch.empty((M, 6), dtype=torch.float64, device=dev)
    cv = torch.empty(M, dtype=torch.float64, device=dev)
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    L = _lib.load()
    _lib.check(L.wnfea_c3d10_stress(M, M * 10, _lib.dptr(conn), _lib.dptr(nodes), _lib.dptr(Et), _lib.dptr(nut),
                                    _lib.dptr(U), None, None, _lib.dptr(gs), _lib.dptr(gv), _lib.dptr(cs),
                                    _lib.dptr(cv), None, _lib.dptr(bad), _lib.stream_of()), "wnfea_c3d10_stress")
    _raise_bad(int(bad.item()), "stresses")
    gs, gv = gs.cpu().numpy(), gv.cpu().numpy()
    if single:
        return gs[0], gv[0], float(gv[0].max())
    return gs, gv, gv.max(axis=1)
