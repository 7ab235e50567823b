"""Block-sparse matrices: host container and the device (gfx950) matrix.

Host side keeps the reference's ``BSR6x6Matrix`` container and conversions
(``wnfea/solver/bsr_matrix.py:19-152``) generalised to a block size b in
{3, 6}: ``block_row_ptr [nb+1] int32``, ``block_col_idx [nnzb] int32``,
``values [nnzb, b*b]`` row-major blocks.  The block pattern of a scalar CSR
is the union of block columns over the b rows of each block row; an empty
block row gets its diagonal block (``bsr_matrix.py:74-88``).

``DeviceBSR`` owns a library matrix handle: the same structure laid out as
sliced ELL on the device (see csrc/sell.hpp), with FP64 values, FP32 values
and optionally FP64 equilibrated values.  SpMV runs on the GPU only.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from .. import _lib


@dataclass
class BSRMatrix:
    block_row_ptr: np.ndarray
    block_col_idx: np.ndarray
    values: np.ndarray
    n_nodes: int
    nnz_blocks: int
    dtype: np.dtype
    block_size: int = 6

    @property
    def shape(self) -> tuple[int, int]:
        n = self.n_nodes * self.block_size
        return (n, n)

    @property
    def nnz(self) -> int:
        return self.nnz_blocks * self.block_size * self.block_size


# the reference's name for the 6x6 container (positional fields identical)
BSR6x6Matrix = BSRMatrix


def csr_to_bsr(K_csr, block_size: int, dtype=np.float64) -> BSRMatrix:
    """Scalar CSR -> block CSR with the reference's pattern rule."""
    K = sp.csr_matrix(K_csr)
    b = int(block_size)
    n_dofs = K.shape[0]
    if n_dofs % b:
        raise ValueError(f"Matrix dimension ({n_dofs}) must be a multiple of {b} for BSR {b}x{b}.")
    n = n_dofs // b
    K.sort_indices()
    rows = np.repeat(np.arange(n_dofs, dtype=np.int64), np.diff(K.indptr))
    br, bc = rows // b, K.indices.astype(np.int64) // b
    keys = np.unique(br * n + bc)
    have = np.zeros(n, dtype=bool)
    have[keys // n] = True
    if not have.all():  # empty block row: diagonal block
        miss = np.flatnonzero(~have).astype(np.int64)
        keys = np.unique(np.concatenate([keys, miss * n + miss]))
    kb, kc = keys // n, keys % n
    rowptr = np.zeros(n + 1, dtype=np.int32)
    rowptr[1:] = np.cumsum(np.bincount(kb, minlength=n))
    vals = np.zeros((len(keys), b * b), dtype=dtype)
    slot = np.searchsorted(keys, br * n + bc)
    vals[slot, (rows % b) * b + (K.indices % b)] = K.data
    return BSRMatrix(rowptr, kc.astype(np.int32), vals, n, len(keys), np.dtype(dtype), b)


def csr_to_bsr6x6(K_csr, dtype=np.float32) -> BSRMatrix:
    return csr_to_bsr(K_csr, 6, dtype)


def csr_to_bsr3x3(K_csr, dtype=np.float64) -> BSRMatrix:
    return csr_to_bsr(K_csr, 3, dtype)


def bsr_to_csr(K_bsr: BSRMatrix, dtype=None) -> sp.csr_matrix:
    """Block CSR -> scalar CSR (sums duplicates), ``bsr_matrix.py:125-152``."""
    b = K_bsr.block_size
    dtype = K_bsr.dtype if dtype is None else dtype
    n = K_bsr.n_nodes
    rr, cc = np.meshgrid(np.arange(b), np.arange(b), indexing="ij")
    rstart = np.repeat(np.arange(n), np.diff(K_bsr.block_row_ptr)) * b
    cstart = np.asarray(K_bsr.block_col_idx, dtype=np.int64) * b
    rows = (rstart[:, None] + rr.ravel()).ravel()
    cols = (cstart[:, None] + cc.ravel()).ravel()
    return sp.csr_matrix((np.asarray(K_bsr.values).ravel().astype(dtype), (rows, cols)), shape=(n * b, n * b))


def bsr6x6_to_csr(K_bsr: BSRMatrix, dtype=None) -> sp.csr_matrix:
    return bsr_to_csr(K_bsr, dtype)


class DeviceBSR:
    """Library-owned sliced-ELL block matrix on the current device."""

    def __init__(self, br: int, bc: int, n_rows: int, n_cols: int, rowptr: np.ndarray, colidx: np.ndarray,
                 sigma: int = 4096, lanes: int = 0):
        import torch
        self.device = _lib.require_gpu()
        self.lib = _lib.load()
        self.br, self.bc = int(br), int(bc)
        self.n_rows, self.n_cols = int(n_rows), int(n_cols)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.colidx = np.ascontiguousarray(colidx, dtype=np.int32)
        if len(self.rowptr) != self.n_rows + 1 or self.rowptr[-1] != len(self.colidx):
            raise ValueError("inconsistent block structure")
        if len(self.colidx) and (self.colidx.min() < 0 or self.colidx.max() >= self.n_cols):
            raise ValueError("block column index out of range")
        self.nnzb = int(self.rowptr[-1])
        self.handle = self.lib.wnfea_bsr_create(self.br, self.bc, self.n_rows, self.n_cols,
                                                _lib.hptr(self.rowptr), _lib.hptr(self.colidx), int(sigma),
                                                int(lanes))
        if not self.handle:
            raise _lib.HipError("wnfea_bsr_create failed")
        rows = np.repeat(np.arange(self.n_rows, dtype=np.int32), np.diff(self.rowptr))
        self.d_rowidx = torch.as_tensor(rows, device=self.device)
        self.d_colidx = torch.as_tensor(self.colidx, device=self.device)

    @classmethod
    def from_host(cls, K: BSRMatrix, which: int = 3, sigma: int = 4096, lanes: int = 0) -> "DeviceBSR":
        import torch
        m = cls(K.block_size, K.block_size, K.n_nodes, K.n_nodes, K.block_row_ptr, K.block_col_idx, sigma, lanes)
        vals = torch.as_tensor(np.ascontiguousarray(K.values, dtype=np.float64), device=m.device)
        m.set_values(vals, which)
        return m

    def close(self):
        if getattr(self, "handle", None):
            self.lib.wnfea_bsr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.n_rows * self.br, self.n_cols * self.bc)

    def info(self) -> dict:
        out = (_lib._ll * 9)()
        _lib.check(self.lib.wnfea_bsr_info(self.handle, out), "wnfea_bsr_info")
        return {"n_slices": out[0], "slots": out[1], "nnzb": out[2], "sigma": out[3], "has_fp64": bool(out[4]),
                "has_fp32": bool(out[5]), "has_fp64_equilibrated": bool(out[6]), "lanes": out[7],
                "has_fp16": bool(out[8])}

    def tune(self, waves: int = 1, xcd_remap: bool = True, unroll: int = 1, precision: str | None = None):
        """SpMV launch shape for every value precision, or for one ("fp64", or
        "fp32" which also covers the FP16 values)."""
        shape = {"waves": int(waves), "xcd_remap": bool(xcd_remap), "unroll": int(unroll)}
        if precision is None:
            _lib.check(self.lib.wnfea_bsr_tune(self.handle, int(waves), int(bool(xcd_remap)), int(unroll)),
                       "wnfea_bsr_tune")
            self.tuning = {"fp64": dict(shape), "fp32": dict(shape)}
            return
        key = "fp64" if precision == "fp64" else "fp32"
        _lib.check(self.lib.wnfea_bsr_tune_precision(self.handle, 0 if key == "fp64" else 1, int(waves),
                                                     int(bool(xcd_remap)), int(unroll)),
                   "wnfea_bsr_tune_precision")
        self.tuning = dict(getattr(self, "tuning", {}) or {}, **{key: shape})

    def set_values(self, vals_std, which: int = 3):
        """Values in standard block layout (nnzb, br*bc) FP64 device tensor; which: 1 FP64, 2 FP32, 3 both."""
        _lib.check(self.lib.wnfea_bsr_set_values(self.handle, _lib.dptr(vals_std), int(which), _lib.stream_of()),
                   "wnfea_bsr_set_values")

    def get_values(self):
        import torch
        out = torch.empty((self.nnzb, self.br * self.bc), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.wnfea_bsr_get_values(self.handle, _lib.dptr(out), _lib.stream_of()),
                   "wnfea_bsr_get_values")
        return out

    def to_host(self) -> BSRMatrix:
        if self.br != self.bc:
            raise ValueError("host container holds square blocks only")
        vals = self.get_values().cpu().numpy()
        return BSRMatrix(self.rowptr.copy(), self.colidx.copy(), vals, self.n_rows, self.nnzb, np.dtype(np.float64),
                         self.br)

    def to_scipy_bsr(self, values=None) -> sp.bsr_matrix:
        vals = (self.get_values() if values is None else values).cpu().numpy()
        return sp.bsr_matrix((vals.reshape(-1, self.br, self.bc), self.colidx, self.rowptr), shape=self.shape)

    def spmv(self, x, y=None, alpha: float = 1.0, beta: float = 0.0, precision: str = "fp64"):
        """y = alpha A x + beta y on the device; precision fp64 | fp32 | fp64_equilibrated."""
        import torch
        prec = {"fp64": 0, "fp32": 1, "fp64_equilibrated": 2}[precision]
        dt = torch.float32 if prec == 1 else torch.float64
        if x.dtype != dt or x.numel() != self.shape[1]:
            raise ValueError("x has the wrong dtype or length")
        if y is None:
            y = torch.zeros(self.shape[0], dtype=dt, device=self.device)
        _lib.check(self.lib.wnfea_bsr_spmv(self.handle, prec, _lib.dptr(x), _lib.dptr(y), float(alpha), float(beta),
                                           _lib.stream_of()), "wnfea_bsr_spmv")
        return y


def bsr_spmv(K_bsr: BSRMatrix, x: np.ndarray, y: np.ndarray | None = None, alpha: float = 1.0,
             beta: float = 0.0) -> np.ndarray:
    """y = alpha K x + beta y for a host block matrix, computed on the GPU.

    FP32 values run the FP32 kernel (FP32 vectors), FP64 values the FP64 one,
    matching the reference's dtype dispatch (``fast_kernels.py:765-832``)."""
    import torch
    fp32 = np.dtype(K_bsr.dtype) == np.float32
    m = DeviceBSR.from_host(K_bsr, which=2 if fp32 else 1)
    dt = torch.float32 if fp32 else torch.float64
    xt = torch.as_tensor(np.ascontiguousarray(x, dtype=np.float32 if fp32 else np.float64), device=m.device)
    yt = None
    if y is not None and beta != 0.0:
        yt = torch.as_tensor(np.ascontiguousarray(y, dtype=np.float32 if fp32 else np.float64), device=m.device)
    out = m.spmv(xt.to(dt), yt, alpha, beta, "fp32" if fp32 else "fp64")
    res = out.cpu().numpy()
    m.close()
    return res


# The reference names of the 6x6 SpMV entry points (both run on the device here).
def hip_spmv_bsr6x6(K_bsr, x, y=None, alpha: float = 1.0, beta: float = 0.0):
    return bsr_spmv(K_bsr, x, y, alpha, beta)


bsr6x6_spmv_cpu = hip_spmv_bsr6x6
