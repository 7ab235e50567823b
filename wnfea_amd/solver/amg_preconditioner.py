"""Smoothed-aggregation AMG for 3-DOF (C3D10) and 6-DOF (beam) systems.

Setup (host, once per matrix) follows the structure of the reference's
``BlockBeamAMGPreconditioner`` (``wnfea/solver/amg_preconditioner.py:147-332``):
symmetric diagonal equilibration D^-1/2 A D^-1/2 (``:180-185``), aggregation,
per-aggregate QR of the rigid-body near-null-space (``:241-281``, generalised
from 6-DOF beam nodes to 3-DOF solid nodes: 3 translations + 3 rotations),
Galerkin coarse operators R A P (``:283-293``), dense coarsest solve.  Unlike
the reference it is a *smoothed* aggregation AMG:

* strength of connection ||A_ij||_F >= theta sqrt(||A_ii||_F ||A_jj||_F) on
  block norms, three-phase aggregation (native, ``wnfea_sa_aggregate``);
* the tentative prolongator is smoothed, P = (I - omega/lambda D^-1 A) P_tent
  with omega = 4/3 by default (the solid solver's hierarchies take 1.6) and
  lambda the power-iteration estimate of rho(D^-1 A);
* the coarse near-null-space is the R factor of each aggregate's QR;
* Chebyshev smoothing on [lambda/ratio, lambda] replaces damped Jacobi.

The V-cycle itself runs on the GPU (``wnfea_amg_vcycle``) in FP32 or FP64;
the coarsest level is a dense FP64 inverse applied on the GPU.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import LinearOperator

from .. import _lib


def rigid_body_modes(coords: np.ndarray, dofs_per_node: int = 3) -> np.ndarray:
    """Near-null-space of 3-D elasticity: (n*b, 6) columns [tx, ty, tz, rx, ry, rz].

    For b = 6 (beam nodes) the rotational DOFs get the unit rotations, as in
    ``amg_preconditioner.py:30-69``."""
    c = np.asarray(coords, dtype=np.float64)
    n, b = len(c), dofs_per_node
    B = np.zeros((n * b, 6))
    x, y, z = c[:, 0], c[:, 1], c[:, 2]
    B[0::b, 0] = 1.0
    B[1::b, 1] = 1.0
    B[2::b, 2] = 1.0
    B[1::b, 3], B[2::b, 3] = -z, y
    B[0::b, 4], B[2::b, 4] = z, -x
    B[0::b, 5], B[1::b, 5] = -y, x
    if b == 6:
        B[3::b, 3] = 1.0
        B[4::b, 4] = 1.0
        B[5::b, 5] = 1.0
    return B


def _bsr_parts(A: sp.bsr_matrix):
    r, c = A.blocksize
    return (np.ascontiguousarray(A.indptr, dtype=np.int32), np.ascontiguousarray(A.indices, dtype=np.int32),
            np.ascontiguousarray(A.data, dtype=np.float64).reshape(-1, r, c))


def _spgemm(A: sp.bsr_matrix, B: sp.bsr_matrix) -> sp.bsr_matrix:
    """A B for block-CSR matrices on host threads (wnfea_host_bsr_spgemm):
    sorted block columns, every block summed in a fixed order."""
    L = _lib.load()
    ra, ca = A.blocksize
    cb2, cb = B.blocksize
    if ca != cb2 or A.shape[1] != B.shape[0]:
        raise ValueError("block shapes of A and B do not chain")
    nr, nc = A.shape[0] // ra, B.shape[1] // cb
    ap, ai, av = _bsr_parts(A)
    bp, bi, bv = _bsr_parts(B)
    cp = np.zeros(nr + 1, dtype=np.int32)
    tot = L.wnfea_host_bsr_spgemm(nr, ra, ca, cb, nc, _lib.hptr(ap), _lib.hptr(ai), _lib.hptr(av), _lib.hptr(bp),
                                  _lib.hptr(bi), _lib.hptr(bv), _lib.hptr(cp), None, None, 0)
    if tot < 0:
        raise _lib.HipError(f"wnfea_host_bsr_spgemm failed ({tot})")
    ci = np.empty(max(tot, 1), dtype=np.int32)
    cv = np.empty((max(tot, 1), ra, cb))
    tot2 = L.wnfea_host_bsr_spgemm(nr, ra, ca, cb, nc, _lib.hptr(ap), _lib.hptr(ai), _lib.hptr(av), _lib.hptr(bp),
                                   _lib.hptr(bi), _lib.hptr(bv), _lib.hptr(cp), _lib.hptr(ci), _lib.hptr(cv), 0)
    if tot2 != tot:
        raise _lib.HipError(f"wnfea_host_bsr_spgemm failed ({tot2})")
    return sp.bsr_matrix((cv[:tot], ci[:tot], cp), shape=(A.shape[0], B.shape[1]))


def _transpose(A: sp.bsr_matrix) -> sp.bsr_matrix:
    """A^T as a block-CSR matrix with transposed blocks (host kernel)."""
    L = _lib.load()
    ra, ca = A.blocksize
    nr, nc = A.shape[0] // ra, A.shape[1] // ca
    ap, ai, av = _bsr_parts(A)
    nnzb = int(ap[-1])
    tp = np.zeros(nc + 1, dtype=np.int32)
    ti = np.empty(max(nnzb, 1), dtype=np.int32)
    tv = np.empty((max(nnzb, 1), ca, ra))
    _lib.check(L.wnfea_host_bsr_transpose(nr, nc, ra, ca, _lib.hptr(ap), _lib.hptr(ai), _lib.hptr(av), _lib.hptr(tp),
                                          _lib.hptr(ti), _lib.hptr(tv)), "wnfea_host_bsr_transpose")
    return sp.bsr_matrix((tv[:nnzb], ti[:nnzb], tp), shape=(A.shape[1], A.shape[0]))


def _symmetrize_galerkin(Ac: sp.bsr_matrix) -> tuple[sp.bsr_matrix, int]:
    """(Ac + Ac^T) / 2 in place, zero diagonal entries set to one (the
    reference's fix of coarse DOFs no fine DOF interpolates,
    amg_preconditioner.py:288-293); returns (Ac, entries fixed)."""
    L = _lib.load()
    b = Ac.blocksize[0]
    ap, ai, av = _bsr_parts(Ac)
    fixed = L.wnfea_host_bsr_symmetrize(Ac.shape[0] // b, b, _lib.hptr(ap), _lib.hptr(ai), _lib.hptr(av), 0)
    if fixed < 0:
        raise _lib.HipError("wnfea_host_bsr_symmetrize: pattern not structurally symmetric")
    return sp.bsr_matrix((av, ai, ap), shape=Ac.shape), int(fixed)


def _matvec(A: sp.spmatrix):
    """x -> A x: the library's threaded host block SpMV for a square BSR with
    blocks of at most 8, else scipy's."""
    if not (sp.isspmatrix_bsr(A) and A.blocksize[0] == A.blocksize[1] and A.blocksize[0] <= 8):
        return lambda x: A @ x
    L = _lib.load()
    b = A.blocksize[0]
    ap, ai, av = _bsr_parts(A)
    n = A.shape[0] // b

    def mv(x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        _lib.check(L.wnfea_host_bsr_spmv(n, b, _lib.hptr(ap), _lib.hptr(ai), _lib.hptr(av), _lib.hptr(x),
                                         _lib.hptr(y), 0), "wnfea_host_bsr_spmv")
        return y
    return mv


def _smooth_prolongator(A: sp.bsr_matrix, P0: sp.bsr_matrix, dinv: np.ndarray, w: float) -> sp.bsr_matrix:
    """P = P0 - w D^-1 A P0 on the pattern of A P0 (which holds P0's blocks:
    every stored node row has its diagonal block)."""
    AP = _spgemm(A, P0)
    b, k = P0.blocksize
    nb, nc = P0.shape[0] // b, P0.shape[1] // k
    rows_ap = np.repeat(np.arange(nb, dtype=np.int64), np.diff(AP.indptr))
    vals = AP.data * ((-w) * dinv.reshape(nb, b)[rows_ap])[:, :, None]
    rows_p0 = np.repeat(np.arange(nb, dtype=np.int64), np.diff(P0.indptr))
    keys_ap = rows_ap * nc + AP.indices
    keys_p0 = rows_p0 * nc + P0.indices
    pos = np.searchsorted(keys_ap, keys_p0)
    if len(pos) and (pos.max() >= len(keys_ap) or not np.array_equal(keys_ap[pos], keys_p0)):
        # a node row without its diagonal block: add P0 in scipy instead
        return (P0 - w * (sp.diags(dinv) @ AP)).tobsr(blocksize=(b, k))
    vals[pos] += P0.data
    return sp.bsr_matrix((vals, AP.indices, AP.indptr), shape=P0.shape)


def _power_lmax(A: sp.spmatrix, dinv: np.ndarray, iters: int = 15, seed: int = 0) -> float:
    """Upper estimate of the largest eigenvalue of D^-1 A (dinv > 0): Lanczos
    on the similar symmetric matrix D^-1/2 A D^-1/2, largest Ritz value plus
    its residual norm.  Lanczos reaches the top of the spectrum in a few
    steps where plain power iteration stalls on clustered eigenvalues; the
    Chebyshev smoother needs a value that is not below the true one."""
    n = A.shape[0]
    mv = _matvec(A)
    dh = np.sqrt(np.abs(dinv))
    v = np.random.default_rng(seed).standard_normal(n)
    v /= np.linalg.norm(v)
    V = [v]
    alpha, beta = [], []
    steps = max(2, min(iters + 5, n))
    for k in range(steps):
        w = dh * mv(dh * V[-1])
        a = float(V[-1] @ w)
        w -= a * V[-1]
        if k:
            w -= beta[-1] * V[-2]
        for q in V:  # full reorthogonalisation (few vectors)
            w -= (q @ w) * q
        alpha.append(a)
        b = float(np.linalg.norm(w))
        if b <= 1e-12 * max(abs(a), 1.0) or k == steps - 1:
            beta.append(b)
            break
        beta.append(b)
        V.append(w / b)
    m = len(alpha)
    T = np.diag(alpha) + np.diag(beta[:m - 1], 1) + np.diag(beta[:m - 1], -1)
    ev, evec = np.linalg.eigh(T)
    theta = float(ev[-1])
    resid = abs(beta[m - 1] * evec[-1, -1])
    return max(theta + resid, 1e-300)


def _strength_graph(A: sp.bsr_matrix, theta: float) -> sp.csr_matrix:
    b = A.blocksize[0]
    n = A.shape[0] // b
    nrm = np.sqrt((A.data.reshape(len(A.data), -1) ** 2).sum(axis=1))
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(A.indptr))
    cols = A.indices.astype(np.int64)
    dn = np.zeros(n)
    on_diag = rows == cols
    dn[rows[on_diag]] = nrm[on_diag]
    strong = (~on_diag) & (nrm > 0.0) & (nrm >= theta * np.sqrt(dn[rows] * dn[cols]))
    G = sp.csr_matrix((np.ones(int(strong.sum()), dtype=np.int8), (rows[strong], cols[strong])), shape=(n, n))
    G.sort_indices()
    return G


@dataclass
class AMGLevelData:
    A: sp.bsr_matrix                 # level operator (equilibrated on level 0)
    dinv: np.ndarray                 # point-Jacobi inverse diagonal
    lmax: float
    P: sp.bsr_matrix | None = None
    R: sp.bsr_matrix | None = None
    n_aggregates: int = 0


def sa_setup(A: sp.bsr_matrix, coords: np.ndarray, fixed_mask: np.ndarray | None = None, theta: float = 0.08,
             theta_decay: float = 1.0, omega: float = 4.0 / 3.0, coarse_size: int = 12000, max_levels: int = 10,
             power_iters: int = 15, lmax_safety: float = 1.05,
             dof_scale: np.ndarray | None = None) -> tuple[list[AMGLevelData], np.ndarray]:
    """Host SA hierarchy of a block matrix.

    When A = S K S is the equilibrated operator (S = diag(1/sqrt|K_ii|)), its
    near-null-space is S^-1 B: pass dof_scale = sqrt|K_ii| so the rigid-body
    modes are scaled accordingly.  Coarsening stops at the first level with
    at most coarse_size rows (default 12 000: a dense inverse of up to 1.1 GB
    in FP64).  Returns (levels, coarse_inverse).  Level l > 0 operators are
    6x6-block."""
    L = _lib.load()
    A = sp.bsr_matrix(A)
    A.sort_indices()
    b = A.blocksize[0]
    if coords is not None:
        B = rigid_body_modes(coords, b)
    else:  # no geometry: the b translation modes only (constant per DOF component)
        B = np.zeros((A.shape[0], b))
        for c in range(b):
            B[c::b, c] = 1.0
    if dof_scale is not None:
        B *= np.asarray(dof_scale, dtype=np.float64)[:, None]
    if fixed_mask is not None:
        B[np.asarray(fixed_mask, dtype=bool)] = 0.0
    levels: list[AMGLevelData] = []
    th = theta
    while True:
        b = A.blocksize[0]
        d = A.diagonal()
        dinv = np.where(np.abs(d) > 0.0, 1.0 / np.where(d == 0.0, 1.0, d), 1.0)
        lmax = _power_lmax(A, dinv, power_iters) * lmax_safety
        if A.shape[0] <= coarse_size or len(levels) + 1 >= max_levels:
            levels.append(AMGLevelData(A=A, dinv=dinv, lmax=lmax))
            break
        G = _strength_graph(A, th)
        n = G.shape[0]
        agg = np.empty(n, dtype=np.int32)
        ptr = np.ascontiguousarray(G.indptr, dtype=np.int32)
        ind = np.ascontiguousarray(G.indices, dtype=np.int32)
        na = int(L.wnfea_sa_aggregate(n, _lib.hptr(ptr), _lib.hptr(ind), _lib.hptr(agg)))
        if na <= 0 or na * 6 >= A.shape[0]:
            levels.append(AMGLevelData(A=A, dinv=dinv, lmax=lmax))
            break
        k = B.shape[1]
        n_blk = int((agg >= 0).sum())
        p_rowptr = np.zeros(n + 1, dtype=np.int32)
        p_colidx = np.zeros(max(n_blk, 1), dtype=np.int32)
        p_vals = np.zeros((max(n_blk, 1), b * k))
        Bc = np.zeros((na * k, k))
        Bc_in = np.ascontiguousarray(B, dtype=np.float64)
        _lib.check(L.wnfea_sa_tentative(n, b, k, _lib.hptr(agg), na, _lib.hptr(Bc_in), _lib.hptr(p_rowptr),
                                        _lib.hptr(p_colidx), _lib.hptr(p_vals), _lib.hptr(Bc)), "wnfea_sa_tentative")
        P0 = sp.bsr_matrix((p_vals[:n_blk].reshape(-1, b, k), p_colidx[:n_blk], p_rowptr), shape=(n * b, na * k))
        # omega: one value, or one per coarsening step (the last repeated)
        om = (omega[min(len(levels), len(omega) - 1)] if isinstance(omega, (list, tuple)) else omega)
        # P = P0 - (omega / lmax) D^-1 A P0;  Ac = R A P with R = P^T, on host
        # threads (wnfea_host_bsr_*), symmetrised
        P = _smooth_prolongator(A, P0, dinv, om / lmax)
        R = _transpose(P)
        Ac = _spgemm(R, _spgemm(A, P))
        # An aggregate that cannot carry all k near-null-space modes (one or
        # two solid nodes, collinear nodes) leaves zero columns in P, hence
        # zero rows and columns in Ac: the symmetrisation gives them a unit
        # diagonal (the reference does the same, amg_preconditioner.py:288-293),
        # so they receive no restricted residual and prolong nothing, and the
        # coarser levels and the dense coarse solve stay nonsingular.
        Ac, _ = _symmetrize_galerkin(Ac)
        levels.append(AMGLevelData(A=A, dinv=dinv, lmax=lmax, P=P, R=R, n_aggregates=na))
        A, B = Ac, Bc
        th *= theta_decay
    # exact coarse solve: the dense inverse of the coarsest operator (SPD),
    # applied as one GEMV per cycle on the device.  A large coarsest level pays:
    # on the I-beam an exact solve on ~11 k rows instead of ~170 rows takes CG
    # from 90 to 50 iterations.
    Acd = levels[-1].A.toarray()
    inv = _coarse_inverse(Acd)
    return levels, inv


def _coarse_inverse(Acd: np.ndarray) -> np.ndarray:
    """Symmetric inverse of the coarsest operator by Cholesky, in FP64 on the
    GPU when one is visible (rocSOLVER through torch.linalg: the 9 762-row
    I-beam level factorises in a fraction of a second there), else on the
    host (scipy).  A matrix that is only semidefinite in floating point gets
    the reference's relative 1e-12 diagonal shift (amg_preconditioner.py:314)
    before a second factorisation; an indefinite one falls back to LU."""
    n = len(Acd)
    if n == 0:
        return np.zeros((0, 0))
    shift = 1e-12 * max(float(np.abs(np.diag(Acd)).max()), 1e-300)
    try:
        import torch
        on_gpu = torch.cuda.is_available()
    except ImportError:
        on_gpu = False
    if on_gpu:
        A = torch.as_tensor(np.ascontiguousarray(Acd, dtype=np.float64), device="cuda")
        inv = None
        for sh in (0.0, shift):
            M = A if sh == 0.0 else A + sh * torch.eye(n, dtype=A.dtype, device=A.device)
            Lc, info = torch.linalg.cholesky_ex(M)
            if int(info.item()) == 0:
                inv = torch.cholesky_inverse(Lc)
                break
        if inv is None:
            inv = torch.linalg.inv(A)
        out = (0.5 * (inv + inv.T)).cpu().numpy()
        del A, inv
        torch.cuda.empty_cache()
        return out
    import scipy.linalg as sla
    for sh in (0.0, shift):
        try:
            c = sla.cho_factor(Acd + sh * np.eye(n) if sh else Acd, lower=True, check_finite=False)
            inv = sla.cho_solve(c, np.eye(n), check_finite=False)
            break
        except np.linalg.LinAlgError:
            continue
    else:
        inv = np.linalg.inv(Acd)
    return 0.5 * (inv + inv.T)


class SmoothedAggregationAMG:
    """GPU V-cycle preconditioner of an equilibrated system (see module doc).

    ``fine`` is the system's DeviceBSR (its FP32 values, and for FP64 cycles
    its FP64 equilibrated values, are the level-0 operator, shared with PCG).
    """

    STORAGE_BITS = {"fp64": 1, "fp32": 2, "fp16": 4}
    # "fp16_fine": FP16 values on the finest level, FP32 on the coarser ones;
    # "fp16_coarse": FP32 on the finest level, FP16 on the coarser ones
    VT_CODE = {"fp64": 0, "fp32": 1, "fp16": 2, "fp16_fine": 3, "fp16_coarse": 4, "fp16_smooth": 5}

    def __init__(self, fine, A_hat: sp.bsr_matrix, coords: np.ndarray, fixed_mask: np.ndarray | None = None,
                 dof_scale: np.ndarray | None = None, theta: float = 0.08, theta_decay: float = 1.0,
                 degree: int = 2, ratio: float = 30.0, coarse_size: int = 12000, max_levels: int = 10,
                 coarse_degree: int | None = None, level_degrees: list | None = None,
                 storage: tuple = ("fp64", "fp32", "fp16"), sigma: int = 4096, waves: int = 1,
                 xcd_remap: bool = True, unroll: int = 1, omega: float = 4.0 / 3.0):
        t0 = time.perf_counter()
        self.lib = _lib.load()
        self.levels, inv = sa_setup(A_hat, coords, fixed_mask, theta=theta, theta_decay=theta_decay, omega=omega,
                                    coarse_size=coarse_size, max_levels=max_levels, dof_scale=dof_scale)
        self.setup_host_sec = time.perf_counter() - t0
        self.degree, self.ratio = degree, ratio
        # Chebyshev degree from level 2 on (the small levels are latency-bound:
        # a lower degree there saves whole kernel launches per cycle)
        self.coarse_degree = degree if coarse_degree is None else coarse_degree
        self.level_degrees = [degree if l < 2 else self.coarse_degree for l in range(len(self.levels))]
        if level_degrees is not None:  # explicit per-level degrees (the last entry repeats)
            self.level_degrees = [int(level_degrees[min(l, len(level_degrees) - 1)]) for l in range(len(self.levels))]
        self.fine = fine
        nl = len(self.levels)
        self.storage = tuple(storage)
        mask = 0
        for name in self.storage:
            mask |= self.STORAGE_BITS[name]
        self.handle = self.lib.wnfea_amg_create(nl, mask)
        if not self.handle:
            raise _lib.HipError("wnfea_amg_create failed")
        self._keep = []
        for l, lv in enumerate(self.levels):
            b = lv.A.blocksize[0]
            n_rows = lv.A.shape[0] // b
            dinv = np.ascontiguousarray(lv.dinv, dtype=np.float64)
            args_A = (None, None, None)
            fine_h = None
            if l == 0:
                fine_h = fine.handle
            else:
                args_A = (np.ascontiguousarray(lv.A.indptr, dtype=np.int32),
                          np.ascontiguousarray(lv.A.indices, dtype=np.int32),
                          np.ascontiguousarray(lv.A.data.reshape(len(lv.A.data), -1), dtype=np.float64))
            if lv.P is not None:
                kc = lv.P.blocksize[1]
                nc = lv.P.shape[1] // kc
                Pa = (np.ascontiguousarray(lv.P.indptr, dtype=np.int32),
                      np.ascontiguousarray(lv.P.indices, dtype=np.int32),
                      np.ascontiguousarray(lv.P.data.reshape(len(lv.P.data), -1), dtype=np.float64))
                Ra = (np.ascontiguousarray(lv.R.indptr, dtype=np.int32),
                      np.ascontiguousarray(lv.R.indices, dtype=np.int32),
                      np.ascontiguousarray(lv.R.data.reshape(len(lv.R.data), -1), dtype=np.float64))
            else:
                kc, nc, Pa, Ra = 0, 0, (None, None, None), (None, None, None)
            self._keep.append((args_A, Pa, Ra, dinv))
            rc = self.lib.wnfea_amg_set_level(
                self.handle, l, fine_h, b, n_rows, _lib.hptr(args_A[0]), _lib.hptr(args_A[1]), _lib.hptr(args_A[2]),
                _lib.hptr(dinv), float(lv.lmax), float(ratio), int(self.level_degrees[l]), int(kc), int(nc),
                _lib.hptr(Pa[0]), _lib.hptr(Pa[1]), _lib.hptr(Pa[2]), _lib.hptr(Ra[0]), _lib.hptr(Ra[1]),
                _lib.hptr(Ra[2]), int(sigma))
            _lib.check(rc, f"wnfea_amg_set_level({l})")
        inv = np.ascontiguousarray(inv, dtype=np.float64)
        self.coarse_inv = inv
        _lib.check(self.lib.wnfea_amg_set_coarse(self.handle, inv.shape[0], _lib.hptr(inv)), "wnfea_amg_set_coarse")
        _lib.check(self.lib.wnfea_amg_finalize(self.handle, int(waves), int(bool(xcd_remap)), int(unroll)),
                   "wnfea_amg_finalize")
        self._keep = None
        self.setup_sec = time.perf_counter() - t0

    def set_tune(self, precision: str, waves: int, unroll: int, xcd_remap: bool):
        """SpMV launch shape of the cycle for "fp64" values or for "fp32" (and
        FP16) values; the constructor's shape applies to both until then."""
        _lib.check(self.lib.wnfea_amg_tune_precision(self.handle, 0 if precision == "fp64" else 1, int(waves),
                                                     int(bool(xcd_remap)), int(unroll)),
                   "wnfea_amg_tune_precision")

    def set_coarse_precision(self, precision: str) -> None:
        """Coarsest solve of the FP32 / FP16 cycles: "fp64" applies the FP64
        dense inverse (needs "fp64" storage), "fp32" (default) its FP32 copy."""
        if precision not in ("fp32", "fp64"):
            raise ValueError(f"unknown coarse precision {precision!r}")
        _lib.check(self.lib.wnfea_amg_set_coarse_fp64(self.handle, int(precision == "fp64")),
                   "wnfea_amg_set_coarse_fp64")

    def stats(self) -> dict:
        out = []
        for lv in self.levels:
            out.append({"n": lv.A.shape[0], "nnz": int(lv.A.nnz), "block": lv.A.blocksize[0],
                        "lmax": lv.lmax, "aggregates": lv.n_aggregates,
                        "P_nnz": int(lv.P.nnz) if lv.P is not None else 0})
        nnz0 = out[0]["nnz"]
        return {"levels": out, "operator_complexity": sum(o["nnz"] for o in out) / max(nnz0, 1),
                "setup_sec": self.setup_sec, "setup_host_sec": self.setup_host_sec}

    def vcycle(self, r, z=None, storage: str | None = None):
        """z = M r on the device.  FP64 vectors run the FP64 cycle; FP32 vectors
        the FP32 cycle, or the FP16-stored one with storage="fp16"."""
        import torch
        if storage is None:
            storage = "fp32" if r.dtype == torch.float32 else "fp64"
        if (storage == "fp64") != (r.dtype == torch.float64):
            raise ValueError("FP64 cycles take FP64 vectors, FP32/FP16 cycles FP32 vectors")
        if z is None:
            z = torch.empty_like(r)
        _lib.check(self.lib.wnfea_amg_vcycle(self.handle, self.VT_CODE[storage], _lib.dptr(r), _lib.dptr(z),
                                             _lib.stream_of()), "wnfea_amg_vcycle")
        return z

    def attach_matrix_free(self, op, scale=None):
        """Run the finest level's smoothing steps and defect of the FP64 and
        FP32 cycles on the matrix-free operator ``op`` (a MatrixFreeOperator
        of the fine system) scaled as the fine level, diag(scale) K
        diag(scale); ``scale`` defaults to the system's equilibration.  None
        detaches.  Attach again after the system's values change (the FP32
        cycle keeps rounded copies).  The hierarchy keeps ``op`` alive while
        attached."""
        prev = getattr(self, "_mf", None)
        if op is not None and scale is None:
            scale = op.system.scale
        _lib.check(self.lib.wnfea_amg_attach_mf(self.handle, op.handle if op is not None else None,
                                                _lib.dptr(scale) if op is not None else None, _lib.stream_of()),
                   "wnfea_amg_attach_mf")
        if prev is not None and prev is not op:
            prev._attached.discard(self)
        self._mf, self._mf_scale = op, scale
        if op is not None:
            op._attached.add(self)
            if getattr(self, "autotuned", False):
                # the cycle's products changed: retune its launch shapes
                from .autotune import autotune_cycle
                for prec in ("fp32", "fp64"):
                    if prec in self.storage:
                        autotune_cycle(self, prec)

    @property
    def matrix_free_smoothing(self) -> bool:
        return getattr(self, "_mf", None) is not None

    def close(self):
        if getattr(self, "handle", None):
            self.lib.wnfea_amg_destroy(self.handle)
            self.handle = None
        mf = getattr(self, "_mf", None)
        if mf is not None:
            mf._attached.discard(self)
            self._mf = self._mf_scale = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _LevelView:
    """Host view of one level in the working precision (A, inv_diag, P, R),
    as the reference's ``AMGLevel`` (``amg_preconditioner.py:128-145``)."""

    def __init__(self, lv, dt):
        self.A = sp.csr_matrix(lv.A).astype(dt)
        self.inv_diag = np.asarray(lv.dinv, dtype=dt)
        self.P = sp.csr_matrix(lv.P).astype(dt) if lv.P is not None else None
        self.R = sp.csr_matrix(lv.R).astype(dt) if lv.R is not None else None


class BlockBeamAMGPreconditioner(LinearOperator):
    """Drop-in for ``wnfea/solver/amg_preconditioner.py:147-661``: M^-1 as a
    scipy LinearOperator (``matvec`` takes and returns FP64 numpy vectors),
    ``solve_pcg``, the tri-precision switch (``enable_tri_precision``,
    ``apply_adaptive``, ``last_precision_used``) and ``levels`` in the working
    precision.  Underneath it is this package's device SA-AMG on the
    symmetrically equilibrated operator (``z = S M(S r)``, as the reference's
    ``inv_sqrt_d`` scaling); every V-cycle runs on the GPU whatever
    ``device`` says (``"cpu"`` is accepted for API parity).  ``A`` (scipy, 6
    or 3 DOF per node) defaults to ``assemble_global_system_sparse(model)``;
    supported DOFs are recognised as unit-diagonal rows without couplings.
    ``coarse_size`` and ``max_levels`` bound the hierarchy as in the
    reference (the coarsest level is solved exactly; the tuned solver path,
    ``solid_solver.build_amg``, uses a 12 000-row coarsest level instead)."""

    _PREC = {0: "fp64", 1: "fp32", 2: "fp16"}

    def __init__(self, model, A=None, max_levels: int = 6, coarse_size: int = 36,
                 working_dtype=np.float64, device: str = "cpu"):
        import torch
        from .solid_solver import _sync, build_amg, system_from_csr
        if A is None:
            from .beam_solver import assemble_global_system_sparse
            A, _ = assemble_global_system_sparse(model)
        A = sp.csr_matrix(A, dtype=np.float64)
        n = A.shape[0]
        super().__init__(dtype=np.float64, shape=(n, n))
        nn = len(model.mesh_nodes)
        dpn = n // nn if nn and n % nn == 0 and n // nn in (3, 6) else 6
        self.model = model
        self.max_levels, self.coarse_size = int(max_levels), int(coarse_size)
        self.working_dtype = np.dtype(working_dtype)
        self.device = str(device).lower()
        self.use_tri_precision = False
        self.switch_tol = 1e-2
        self.last_precision_used = None
        self.fp16_vcycle_count = 0
        self.fp32_vcycle_count = 0
        # supported DOFs: a unit diagonal and nothing else in the row
        nnz = np.diff(A.indptr)
        fixed = (nnz == 1) & (A.diagonal() == 1.0)
        self.system = system_from_csr(A, np.zeros(n), fixed_mask=fixed,
                                      coords=np.asarray(model.mesh_nodes, dtype=np.float64), dpn=dpn)
        S = self.system.scale.cpu().numpy()
        self.inv_sqrt_d = S
        self.amg = build_amg(self.system, storage=("fp64", "fp32", "fp16"), coarse_size=self.coarse_size,
                             max_levels=self.max_levels)
        self.gpu_handle = self.amg.handle
        self._A = A
        self._levels = None
        self._S = torch.as_tensor(S, device=self.system.K.device)
        _sync()

    @property
    def levels(self):
        if self._levels is None:
            self._levels = [_LevelView(lv, self.working_dtype) for lv in self.amg.levels]
        return self._levels

    def enable_tri_precision(self, switch_tol: float = 1e-2) -> None:
        """FP16 V-cycles while the relative residual is above switch_tol
        (``amg_preconditioner.py:442-445``)."""
        self.use_tri_precision = True
        self.switch_tol = float(switch_tol)

    def _apply(self, r, storage: str) -> np.ndarray:
        import torch
        rt = torch.as_tensor(np.ascontiguousarray(r, dtype=np.float64), device=self._S.device) * self._S
        if storage != "fp64":
            rt = rt.float()
        z = self.amg.vcycle(rt, storage=storage).double() * self._S
        return z.cpu().numpy()

    def _matvec(self, r):
        storage = "fp32" if self.working_dtype == np.float32 else "fp64"
        self.last_precision_used = self.working_dtype.name
        if storage == "fp32":
            self.fp32_vcycle_count += 1
        return self._apply(np.ravel(r), storage)

    def _rmatvec(self, r):
        return self._matvec(r)  # the FP64/FP32 cycles are symmetric

    def apply_adaptive(self, r, current_res_rel: float | None = None) -> np.ndarray:
        """``amg_preconditioner.py:538-584``: an FP16-stored cycle above the
        switch tolerance, the working precision below it."""
        if self.use_tri_precision and current_res_rel is not None and current_res_rel > self.switch_tol:
            self.last_precision_used = "float16"
            self.fp16_vcycle_count += 1
            return self._apply(np.ravel(r), "fp16")
        return self._matvec(r)

    def solve_pcg(self, b, rtol: float = 1e-6, max_iter: int = 500, precision_mode: int | None = None,
                  check_interval: int = 1, coarse_sweeps: int | None = None):
        """``amg_preconditioner.py:608-661``: PCG on the device with the
        Krylov vectors, the products with A and alpha/beta in FP64 and the
        V-cycle in precision_mode (0 FP64, 1 FP32, 2 FP16-stored; default: the
        working precision).  The coarsest level is solved exactly, so
        coarse_sweeps has no effect.  Returns (u, iterations, ||b - A u|| / ||b||)."""
        import torch
        from .solid_solver import pcg_device
        if precision_mode is None:
            precision_mode = 1 if self.working_dtype == np.float32 else 0
        bt = torch.as_tensor(np.ascontiguousarray(b, dtype=np.float64), device=self._S.device)
        x, it, _, _ = pcg_device(self.system, b=bt, rtol=rtol, max_iter=max_iter, preconditioner="amg",
                                 precision="fp64", amg=self.amg, amg_precision=self._PREC[int(precision_mode)],
                                 check_every=max(1, int(check_interval)))
        u = x.cpu().numpy()
        nb = float(np.linalg.norm(b))
        rel = float(np.linalg.norm(b - self._A @ u) / nb) if nb > 0 else 0.0
        return u, it, rel

    def close(self):
        if getattr(self, "amg", None) is not None:
            self.amg.close()
            self.amg = None
            self.gpu_handle = None
