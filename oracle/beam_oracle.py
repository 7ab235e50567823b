"""numpy restatement of the reference's co-rotational beam path -- TEST
INFRASTRUCTURE ONLY (imported by tests/ as the checker of the HIP beam
kernels; the product never calls it).

* ``corot_forces``      -- vectorised internal forces, wnfea/solver/fast_kernels.py:451-537
                           with the reference triads of fast_kernels.py:297-305
* ``corot_element``     -- the single-element function of
                           wnfea/solver/corotational_beam.py:15-233 (its own triad
                           rule, Rodrigues, twist, B^T q), for cross-checking the
                           two reference formulations
* ``ke_linear``         -- build_element_stiffness_3d_beam + build_transformation_matrix,
                           wnfea/solver/assembler.py:13-148
* ``fd_tangent``        -- central differences of ``corot_forces`` per element
* ``assemble``          -- scalar 6-DOF CSR of element matrices with the Dirichlet
                           rule (zero row/column, unit diagonal)
* ``newton``            -- load-stepped Newton with the assembled symmetric FD
                           tangent and exact sparse solves (the reference's
                           solve_nonlinear_jfnk loop, nonlinear_solver.py:159-376,
                           with a direct inner solve instead of JFNK-PCG)

Pins (tests/test_cpu_beam.py): the two reference formulations agree to
round-off on circular sections; rigid translations and finite rigid rotations
give zero force; the small-displacement limit is the linear stiffness (for
right-handed frames); the FD tangent matches directional derivatives.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def ref_triads(X1, X2):
    v0 = X2 - X1
    L0 = np.linalg.norm(v0, axis=1)
    e1 = v0 / np.maximum(L0[:, None], 1e-15)
    ref = np.where((np.abs(e1[:, 2]) < 0.9)[:, None], np.array([0.0, 0.0, 1.0]), np.array([0.0, 1.0, 0.0]))
    e2 = _unit(np.cross(ref, e1))
    e3 = _unit(np.cross(e1, e2))
    return e1, e2, e3, L0


def corot_element_forces(X1, X2, props, ue):
    """Per-element f (M, 12) for element displacements ue (M, 12)."""
    e1, e2, e3, L0 = ref_triads(X1, X2)
    v0 = X2 - X1
    u1, th1, u2, th2 = ue[:, 0:3], ue[:, 3:6], ue[:, 6:9], ue[:, 9:12]
    du = u2 - u1
    v = v0 + du
    L = np.linalg.norm(v, axis=1)
    r1 = v / np.maximum(L[:, None], 1e-15)
    cr = np.cross(e1, r1)
    sphi = np.linalg.norm(cr, axis=1)
    cphi = np.sum(e1 * r1, axis=1)
    r2, r3 = e2.copy(), e3.copy()
    m = sphi > 1e-12
    if np.any(m):
        ax = cr[m] / sphi[m, None]
        phi = np.arctan2(sphi[m], cphi[m])
        cp, sn = np.cos(phi)[:, None], np.sin(phi)[:, None]
        r2[m] = e2[m] * cp + np.cross(ax, e2[m]) * sn + ax * np.sum(ax * e2[m], axis=1, keepdims=True) * (1 - cp)
        r3[m] = e3[m] * cp + np.cross(ax, e3[m]) * sn + ax * np.sum(ax * e3[m], axis=1, keepdims=True) * (1 - cp)
    tw = 0.5 * np.sum((th1 + th2) * r1, axis=1)
    ct, st = np.cos(tw)[:, None], np.sin(tw)[:, None]
    r2t = r2 * ct + np.cross(r1, r2) * st
    r1 = _unit(r1)
    r2 = _unit(r2t - np.sum(r2t * r1, axis=1, keepdims=True) * r1)
    r3 = _unit(np.cross(r1, r2))
    ul = (2.0 * np.sum(v0 * du, axis=1) + np.sum(du * du, axis=1)) / (L + L0)
    thy1 = np.sum(r2 * th1, axis=1) + np.sum(r3 * du, axis=1) / L
    thz1 = np.sum(r3 * th1, axis=1) - np.sum(r2 * du, axis=1) / L
    thx = np.sum(r1 * (th2 - th1), axis=1)
    thy2 = np.sum(r2 * th2, axis=1) + np.sum(r3 * du, axis=1) / L
    thz2 = np.sum(r3 * th2, axis=1) - np.sum(r2 * du, axis=1) / L
    E, G, A, Iy, Iz, J = (props[:, k] for k in range(6))
    N = E * A / L0 * ul
    My1 = E * Iy / L0 * (4 * thy1 + 2 * thy2)
    My2 = E * Iy / L0 * (2 * thy1 + 4 * thy2)
    Mz1 = E * Iz / L0 * (4 * thz1 + 2 * thz2)
    Mz2 = E * Iz / L0 * (2 * thz1 + 4 * thz2)
    Tx = G * J / L0 * thx
    f1 = -N[:, None] * r1 - ((My1 + My2) / L)[:, None] * r3 + ((Mz1 + Mz2) / L)[:, None] * r2
    m1 = -Tx[:, None] * r1 + My1[:, None] * r2 + Mz1[:, None] * r3
    m2 = Tx[:, None] * r1 + My2[:, None] * r2 + Mz2[:, None] * r3
    return np.hstack([f1, m1, -f1, m2])


def element_dofs(conn):
    conn = np.asarray(conn, dtype=np.int64)
    return np.hstack([6 * conn[:, :1] + np.arange(6), 6 * conn[:, 1:2] + np.arange(6)])


def corot_forces(nodes, conn, props, U):
    """Global F_int (N*6) = sum over beams of B^T q (fast_kernels.py:451-537)."""
    conn = np.asarray(conn, dtype=np.int64)
    dofs = element_dofs(conn)
    fe = corot_element_forces(nodes[conn[:, 0]], nodes[conn[:, 1]], np.asarray(props, dtype=np.float64),
                              np.asarray(U, dtype=np.float64)[dofs])
    F = np.zeros(6 * len(nodes))
    np.add.at(F, dofs.ravel(), fe.ravel())
    return F


def corot_element(node1, node2, u_e, E, G, A, Iy, Iz, J):
    """The single-element function of corotational_beam.py:15-233 (its own
    triad rule and the explicit 6x12 projector)."""
    X1, X2 = np.asarray(node1, float), np.asarray(node2, float)
    v0 = X2 - X1
    L0 = np.linalg.norm(v0)
    e1 = v0 / L0
    if np.isclose(abs(e1[0]), 1.0):
        e2, e3 = np.array([0.0, 1.0, 0.0]), np.array([0.0, 0.0, 1.0])
    elif np.isclose(abs(e1[1]), 1.0):
        e2, e3 = np.array([1.0, 0.0, 0.0]), np.array([0.0, 0.0, 1.0])
    elif np.isclose(abs(e1[2]), 1.0):
        e2, e3 = np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0])
    else:
        e2 = np.cross([0.0, 0.0, 1.0], e1)
        if np.linalg.norm(e2) < 1e-9:
            e2 = np.cross([0.0, 1.0, 0.0], e1)
        e2 /= np.linalg.norm(e2)
        e3 = np.cross(e1, e2)
        e3 /= np.linalg.norm(e3)
    u1, t1, u2, t2 = u_e[0:3], u_e[3:6], u_e[6:9], u_e[9:12]
    if np.linalg.norm(u_e) < 1e-15:
        return np.zeros(12)
    v = v0 + (u2 - u1)
    L = np.linalg.norm(v)
    r1 = v / L
    cr = np.cross(e1, r1)
    s, c = np.linalg.norm(cr), float(e1 @ r1)
    if s > 1e-12:
        ax, phi = cr / s, np.arctan2(s, c)

        def rot(w):
            return w * np.cos(phi) + np.cross(ax, w) * np.sin(phi) + ax * (ax @ w) * (1 - np.cos(phi))
        r2, r3 = rot(e2), rot(e3)
    elif c < 0:
        r2, r3 = -e2, -e3
    else:
        r2, r3 = e2.copy(), e3.copy()
    tw = 0.5 * float((t1 + t2) @ r1)
    if abs(tw) > 1e-12:
        r2, r3 = r2 * np.cos(tw) + np.cross(r1, r2) * np.sin(tw), r3 * np.cos(tw) + np.cross(r1, r3) * np.sin(tw)
    r2 = r2 - (r2 @ r1) * r1
    r2 /= np.linalg.norm(r2)
    r3 = np.cross(r1, r2)
    r3 /= np.linalg.norm(r3)
    du = u2 - u1
    ul = (2 * v0 @ du + du @ du) / (L + L0)
    thy1, thz1 = r2 @ t1 + r3 @ du / L, r3 @ t1 - r2 @ du / L
    thx = r1 @ (t2 - t1)
    thy2, thz2 = r2 @ t2 + r3 @ du / L, r3 @ t2 - r2 @ du / L
    q = np.array([E * A / L0 * ul, E * Iy / L0 * (4 * thy1 + 2 * thy2), E * Iz / L0 * (4 * thz1 + 2 * thz2),
                  G * J / L0 * thx, E * Iy / L0 * (2 * thy1 + 4 * thy2), E * Iz / L0 * (2 * thz1 + 4 * thz2)])
    B = np.zeros((6, 12))
    B[0, 0:3], B[0, 6:9] = -r1, r1
    B[1, 0:3], B[1, 3:6], B[1, 6:9] = -r3 / L, r2, r3 / L
    B[2, 0:3], B[2, 3:6], B[2, 6:9] = r2 / L, r3, -r2 / L
    B[3, 3:6], B[3, 9:12] = -r1, r1
    B[4, 0:3], B[4, 6:9], B[4, 9:12] = -r3 / L, r3 / L, r2
    B[5, 0:3], B[5, 6:9], B[5, 9:12] = r2 / L, -r2 / L, r3
    return B.T @ q


def ke_local(L, E, G, A, Iy, Iz, J):
    K = np.zeros((12, 12))
    ax, tor = E * A / L, G * J / L
    K[0, 0] = K[6, 6] = ax
    K[0, 6] = K[6, 0] = -ax
    K[3, 3] = K[9, 9] = tor
    K[3, 9] = K[9, 3] = -tor
    b1, b2, b3, b4 = 12 * E * Iz / L ** 3, 6 * E * Iz / L ** 2, 4 * E * Iz / L, 2 * E * Iz / L
    K[1, 1] = K[7, 7] = b1
    K[1, 7] = K[7, 1] = -b1
    K[1, 5] = K[5, 1] = K[1, 11] = K[11, 1] = b2
    K[5, 5] = K[11, 11] = b3
    K[5, 11] = K[11, 5] = b4
    K[5, 7] = K[7, 5] = K[7, 11] = K[11, 7] = -b2
    c1, c2, c3, c4 = 12 * E * Iy / L ** 3, 6 * E * Iy / L ** 2, 4 * E * Iy / L, 2 * E * Iy / L
    K[2, 2] = K[8, 8] = c1
    K[2, 8] = K[8, 2] = -c1
    K[2, 4] = K[4, 2] = K[2, 10] = K[10, 2] = -c2
    K[4, 4] = K[10, 10] = c3
    K[4, 10] = K[10, 4] = c4
    K[4, 8] = K[8, 4] = K[8, 10] = K[10, 8] = c2
    return K


def frame(X1, X2):
    v = X2 - X1
    L = np.linalg.norm(v)
    if np.isclose(abs(v[0]), L):
        return np.array([[1.0 if v[0] > 0 else -1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]])
    if np.isclose(abs(v[1]), L):
        return np.array([[0, 1.0 if v[1] > 0 else -1.0, 0], [1.0, 0, 0], [0, 0, 1.0]])
    if np.isclose(abs(v[2]), L):
        return np.array([[0, 0, 1.0 if v[2] > 0 else -1.0], [1.0, 0, 0], [0, 1.0, 0]])
    lx = v / L
    ly = np.cross([0.0, 0.0, 1.0], lx)
    if np.linalg.norm(ly) < 1e-9:
        ly = np.cross([0.0, 1.0, 0.0], lx)
    ly /= np.linalg.norm(ly)
    lz = np.cross(lx, ly)
    return np.array([lx, ly, lz / np.linalg.norm(lz)])


def ke_linear(nodes, conn, props):
    conn = np.asarray(conn, dtype=np.int64)
    out = np.zeros((len(conn), 12, 12))
    for e, (a, b) in enumerate(conn):
        X1, X2 = nodes[a], nodes[b]
        L = np.linalg.norm(X2 - X1)
        if L < 1e-12:
            continue
        T = np.kron(np.eye(4), frame(X1, X2))
        out[e] = T.T @ ke_local(L, *props[e]) @ T
    return out


def fd_tangent(nodes, conn, props, U, h_rel=1e-6):
    """Kt[e] = d f_e / d u_e by central differences (h = h_rel L0 for
    translations, h_rel for rotations)."""
    conn = np.asarray(conn, dtype=np.int64)
    X1, X2 = nodes[conn[:, 0]], nodes[conn[:, 1]]
    ue = np.asarray(U, dtype=np.float64)[element_dofs(conn)]
    L0 = np.linalg.norm(X2 - X1, axis=1)
    Kt = np.zeros((len(conn), 12, 12))
    for j in range(12):
        h = h_rel * L0 if (j % 6) < 3 else np.full(len(conn), h_rel)
        up, um = ue.copy(), ue.copy()
        up[:, j] += h
        um[:, j] -= h
        Kt[:, :, j] = (corot_element_forces(X1, X2, props, up) - corot_element_forces(X1, X2, props, um)) / (2 * h[:, None])
    return Kt


def assemble(n_nodes, conn, Ke, fixed_mask, sym=True):
    dofs = element_dofs(conn)
    K = Ke if not sym else 0.5 * (Ke + np.transpose(Ke, (0, 2, 1)))
    r = np.repeat(dofs, 12, axis=1).ravel()
    c = np.tile(dofs, (1, 12)).ravel()
    n = 6 * n_nodes
    A = sp.coo_matrix((K.ravel(), (r, c)), shape=(n, n)).tocsr()
    fixed = np.asarray(fixed_mask, dtype=bool)
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    A.data[(fixed[rows] | fixed[A.indices]) & (rows != A.indices)] = 0.0
    A.data[fixed[rows] & (rows == A.indices)] = 1.0
    A.eliminate_zeros()
    return A


def newton(nodes, conn, props, F_ext, fixed_mask, n_steps=5, tol_rel=1e-8, max_iter=30):
    """Load-stepped Newton with the symmetric FD tangent and exact solves.
    Returns U (N*6) and the Newton iterations per step."""
    fixed = np.asarray(fixed_mask, dtype=bool)
    U = np.zeros(6 * len(nodes))
    its = []
    for lam in np.linspace(1.0 / n_steps, 1.0, n_steps):
        R = corot_forces(nodes, conn, props, U) - lam * F_ext
        R[fixed] = U[fixed]
        r0 = np.linalg.norm(R)
        k = 0
        while np.linalg.norm(R) > tol_rel * max(r0, 1e-30) and k < max_iter:
            Kt = assemble(len(nodes), conn, fd_tangent(nodes, conn, props, U), fixed)
            U = U + spla.spsolve(Kt.tocsc(), -R)
            R = corot_forces(nodes, conn, props, U) - lam * F_ext
            R[fixed] = U[fixed]
            k += 1
        its.append(k)
    return U, its
