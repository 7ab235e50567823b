"""numpy glue over the C oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, as the checker of the HIP library.  It
restates the reference's solve path on the CPU in FP64:

* ``assemble_csr``  -- ``wnfea/solver/solid_solver.py:26-129`` (scalar CSR,
  element-order duplicate summation, BC row/col zeroing, unit diagonal,
  F = 0 at fixed DOFs, ``eliminate_zeros``);
* ``csr_to_bsr``     -- ``wnfea/solver/bsr_matrix.py:49-122`` generalised to
  block size b (union of block columns of the b rows, diagonal block for an
  empty block row);
* ``rcm``            -- scipy's reverse Cuthill-McKee (``wnfea/mesh/rcm.py:85``)
  with numpy's argsort of the degrees as the seed order;
* element, PCG, mixed IR and stress functions -- see wnfea_oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_p = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-C", _HERE, os.path.join(_HERE, "liboracle.so")], check=True,
                           stdout=subprocess.DEVNULL)
        L = ctypes.CDLL(path)
        L.or_c3d10_ke.argtypes = [_p, _d, _d, _p]
        L.or_c3d10_fint.argtypes = [_p, _p, _d, _d, _p]
        L.or_c3d10_stress.argtypes = [_p, _p, _d, _d, _p, _p]
        L.or_ke_batch.argtypes = [_i, _p, _p, _p, _p, _p]
        L.or_fint_global_stride.argtypes = [_i, _p, _p, _p, _p, _p, _i]
        L.or_csr_accumulate.argtypes = [_i, _p, _p, _p, _p, _p]
        L.or_csr_spmv.argtypes = [_i, _p, _p, _p, _p, _p]
        L.or_pcg.argtypes = [_i, _p, _p, _p, _p, _p, _d, _i, _i, ctypes.POINTER(_d)]
        L.or_ir_mixed.argtypes = [_i, _p, _p, _p, _p, _p, _d, _i, ctypes.POINTER(_d)]
        L.or_rcm.argtypes = [_i, _p, _p, _p, _p]
        L.or_stress_field.argtypes = [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p]
        L.or_shape_derivs.argtypes = [_d, _d, _d, _p, _p]
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own kernels.cpp built from source into oracle/_ref (None if absent)."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_kernels.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.compute_c3d10_internal_forces_native.argtypes = [_p, _p, _p, _p, _p, _i, _i]
        L.compute_c3d10_internal_forces_native.restype = None
        _REF = L
    return _REF


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------------------------------------------------------- element
def shape_derivs(xi, eta, zeta):
    N = np.zeros(10)
    dN = np.zeros((3, 10))
    lib().or_shape_derivs(xi, eta, zeta, _ptr(N), _ptr(dN))
    return N, dN


def element_stiffness(coords, E, nu):
    X = _c(coords, np.float64)
    Ke = np.zeros((30, 30))
    if lib().or_c3d10_ke(_ptr(X), float(E), float(nu), _ptr(Ke)):
        raise ValueError("Non-positive Jacobian determinant in C3D10 element")
    return Ke


def element_internal_forces(coords, ue, E, nu):
    X, u = _c(coords, np.float64), _c(ue, np.float64)
    f = np.zeros(30)
    if lib().or_c3d10_fint(_ptr(X), _ptr(u), float(E), float(nu), _ptr(f)):
        raise ValueError("Non-positive Jacobian determinant in C3D10 element")
    return f


def element_stresses(coords, ue, E, nu):
    X, u = _c(coords, np.float64), _c(ue, np.float64)
    s = np.zeros((4, 6))
    vm = np.zeros(4)
    if lib().or_c3d10_stress(_ptr(X), _ptr(u), float(E), float(nu), _ptr(s), _ptr(vm)):
        raise ValueError("Non-positive Jacobian determinant in C3D10 element")
    return s, vm, float(vm.max())


def ke_batch(nodes, conn, E, nu):
    nodes, conn = _c(nodes, np.float64), _c(conn, np.int32)
    M = len(conn)
    E = _c(np.broadcast_to(E, (M,)), np.float64)
    nu = _c(np.broadcast_to(nu, (M,)), np.float64)
    Ke = np.zeros((M, 30, 30))
    bad = lib().or_ke_batch(M, _ptr(conn), _ptr(nodes), _ptr(E), _ptr(nu), _ptr(Ke))
    if bad:
        raise ValueError(f"Non-positive Jacobian determinant in element {bad - 1}")
    return Ke


def fint_global(nodes, conn, props, U, stride=6):
    nodes, conn, props, U = _c(nodes, np.float64), _c(conn, np.int32), _c(props, np.float64), _c(U, np.float64)
    F = np.zeros(len(nodes) * stride)
    lib().or_fint_global_stride(len(conn), _ptr(conn), _ptr(nodes), _ptr(props), _ptr(U), _ptr(F), stride)
    return F


# ---------------------------------------------------------------- assembly
def node_pattern(n_nodes, conn):
    """Node-level sparsity (union of 10x10 element cliques), sorted CSR, int8 ones."""
    el = np.asarray(conn, dtype=np.int64)
    r = np.repeat(el, 10, axis=1).ravel()
    c = np.tile(el, (1, 10)).ravel()
    A = sp.csr_matrix((np.ones(len(r), dtype=np.int8), (r, c)), shape=(n_nodes, n_nodes))
    A.sum_duplicates()
    A.data[:] = 1
    return A


def assemble_csr(nodes, conn, E, nu, fixed_mask=None, F=None, apply_bcs=True):
    """Scalar 3-DOF CSR stiffness with the reference BC semantics (solid_solver.py:26-129)."""
    n = len(nodes)
    conn = _c(conn, np.int32)
    Ke = ke_batch(nodes, conn, E, nu)
    Pn = node_pattern(n, conn)
    # scalar pattern: each node entry expands to a 3x3 block, columns sorted
    rows_n = np.repeat(np.arange(n), np.diff(Pn.indptr))
    cols_n = Pn.indices.astype(np.int64)
    deg = np.diff(Pn.indptr)
    indptr = np.zeros(3 * n + 1, dtype=np.int64)
    indptr[1:] = np.cumsum(np.repeat(3 * deg, 3))
    # row 3i+r has columns 3*c_k + (0,1,2) for c_k in the node's sorted neighbours
    idx = np.empty(indptr[-1], dtype=np.int64)
    for r in range(3):
        # positions of row 3i+r entries
        starts = indptr[3 * np.arange(n) + r]
        off = np.repeat(starts - 3 * Pn.indptr[:-1], np.diff(Pn.indptr))
        base = 3 * np.arange(len(cols_n)) + off
        for c in range(3):
            idx[base + c] = 3 * cols_n + c
    del rows_n
    indptr32, idx32 = indptr.astype(np.int32), idx.astype(np.int32)
    data = np.zeros(len(idx32))
    if lib().or_csr_accumulate(len(conn), _ptr(conn), _ptr(Ke), _ptr(indptr32), _ptr(idx32), _ptr(data)):
        raise RuntimeError("pattern does not cover an element entry")
    K = sp.csr_matrix((data, idx32, indptr32), shape=(3 * n, 3 * n))
    F = np.zeros(3 * n) if F is None else np.array(F, dtype=np.float64)
    if apply_bcs and fixed_mask is not None and np.any(fixed_mask):
        fixed = np.asarray(fixed_mask, dtype=bool)
        rows = np.repeat(np.arange(3 * n), np.diff(K.indptr))
        mr, mc = fixed[rows], fixed[K.indices]
        K.data[(mr | mc) & (rows != K.indices)] = 0.0
        K.data[mr & (rows == K.indices)] = 1.0
        F[fixed] = 0.0
        K.eliminate_zeros()
    return K, F


def csr_to_bsr(K, b):
    """Block CSR of block size b (bsr_matrix.py:49-122 rule, generalised).
    Returns (block_row_ptr int32, block_col_idx int32, values (nnzb, b*b))."""
    K = sp.csr_matrix(K)
    K.sort_indices()
    n = K.shape[0] // b
    rows = np.repeat(np.arange(K.shape[0]), np.diff(K.indptr))
    br, bc = rows // b, K.indices // b
    keys = np.unique(br.astype(np.int64) * n + bc)
    kb, kc = keys // n, keys % n
    empty = np.setdiff1d(np.arange(n), kb)
    if len(empty):
        keys = np.unique(np.concatenate([keys, empty.astype(np.int64) * n + empty]))
        kb, kc = keys // n, keys % n
    rowptr = np.zeros(n + 1, dtype=np.int32)
    rowptr[1:] = np.cumsum(np.bincount(kb, minlength=n))
    vals = np.zeros((len(keys), b * b), dtype=K.dtype)
    slot = np.searchsorted(keys, br.astype(np.int64) * n + bc)
    vals[slot, (rows % b) * b + (K.indices % b)] = K.data
    return rowptr, kc.astype(np.int32), vals


def bsr_spmv(rowptr, colidx, vals, x, b):
    """y = K x for a block CSR (bsr_matrix.py:155-193 semantics), FP64."""
    n = len(rowptr) - 1
    blk = vals.reshape(-1, b, b).astype(np.float64)
    xb = np.asarray(x, dtype=np.float64).reshape(-1, b)
    contrib = np.einsum("kij,kj->ki", blk, xb[colidx])
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    y = np.zeros((n, b))
    np.add.at(y, rows, contrib)
    return y.ravel()


# ---------------------------------------------------------------- solvers
def _csr_args(K):
    K = sp.csr_matrix(K)
    return (K.shape[0], _c(K.indptr, np.int32), _c(K.indices, np.int32), _c(K.data, np.float64))


def csr_spmv(K, x):
    n, ip, ix, d = _csr_args(K)
    x = _c(x, np.float64)
    y = np.zeros(n)
    lib().or_csr_spmv(n, _ptr(ip), _ptr(ix), _ptr(d), _ptr(x), _ptr(y))
    return y


def pcg(K, F, rtol=1e-6, max_iter=2000, jacobi=True):
    n, ip, ix, d = _csr_args(K)
    b = _c(F, np.float64)
    x = np.zeros(n)
    rel = _d(0.0)
    it = lib().or_pcg(n, _ptr(ip), _ptr(ix), _ptr(d), _ptr(b), _ptr(x), rtol, max_iter, int(jacobi), ctypes.byref(rel))
    return x, it, rel.value


def ir_mixed(K, F, rtol=1e-6, max_iter=2000):
    n, ip, ix, d = _csr_args(K)
    b = _c(F, np.float64)
    x = np.zeros(n)
    rel = _d(0.0)
    it = lib().or_ir_mixed(n, _ptr(ip), _ptr(ix), _ptr(d), _ptr(b), _ptr(x), rtol, max_iter, ctypes.byref(rel))
    return x, it, rel.value


def rcm(adj, seed_order=None):
    """RCM order new_to_old of a symmetric CSR graph (scipy semantics)."""
    A = sp.csr_matrix(adj)
    A.sort_indices()
    n = A.shape[0]
    ip, ix = _c(A.indptr, np.int32), _c(A.indices, np.int32)
    if seed_order is None:
        deg = np.diff(A.indptr).astype(np.int32)
        has_diag = np.zeros(n, dtype=np.int32)
        rows = np.repeat(np.arange(n), np.diff(A.indptr))
        has_diag[rows[rows == A.indices]] = 1
        seed_order = np.argsort(deg + has_diag)
    so = _c(seed_order, np.int32)
    perm = np.zeros(n, dtype=np.int32)
    lib().or_rcm(n, _ptr(ip), _ptr(ix), _ptr(so), _ptr(perm))
    return perm


def stress_field(nodes, conn, E, nu, u3):
    nodes, conn, u3 = _c(nodes, np.float64), _c(conn, np.int32), _c(u3, np.float64)
    M, n = len(conn), len(nodes)
    E = _c(np.broadcast_to(E, (M,)), np.float64)
    nu = _c(np.broadcast_to(nu, (M,)), np.float64)
    cs, cv, nv = np.zeros((M, 6)), np.zeros(M), np.zeros(n)
    bad = lib().or_stress_field(n, M, _ptr(conn), _ptr(nodes), _ptr(E), _ptr(nu), _ptr(u3), _ptr(cs), _ptr(cv), _ptr(nv))
    if bad:
        raise ValueError(f"Non-positive Jacobian determinant in element {bad - 1}")
    return cs, cv, nv


# ---------------------------------------------------------------- sensitivities
def adjoint_grad(nodes, conn, E, nu, u3, lam3, chunk=20000):
    """g_e = -lam_e^T (K_e / E_e) u_e with K_e from the oracle element routine
    (K is linear in E_e, so dK/dE_e = K_e / E_e).  The reference has no
    sensitivities; this is the closed form the device gradient is held to."""
    conn = _c(conn, np.int32)
    M = len(conn)
    E = np.broadcast_to(np.asarray(E, dtype=np.float64), (M,))
    nu = np.broadcast_to(np.asarray(nu, dtype=np.float64), (M,))
    u = np.asarray(u3, dtype=np.float64).reshape(-1, 3)
    lam = np.asarray(lam3, dtype=np.float64).reshape(-1, 3)
    g = np.empty(M)
    for a in range(0, M, chunk):
        b = min(M, a + chunk)
        Ke = ke_batch(nodes, conn[a:b], E[a:b], nu[a:b])
        ue = u[conn[a:b]].reshape(b - a, 30)
        le = lam[conn[a:b]].reshape(b - a, 30)
        g[a:b] = -np.einsum("mi,mij,mj->m", le, Ke, ue) / E[a:b]
    return g

