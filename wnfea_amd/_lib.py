"""ctypes bridge to libwnfea_hip.so (the gfx950 C-ABI declared in include/wnfea_hip.h).

The library is loaded after ``torch`` so that it binds to the HIP runtime
torch already loaded (one runtime per process; device pointers from torch
tensors are then valid in the library).  There is no CPU fallback: if the
library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwnfea_hip.so")

_p = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_ll = ctypes.c_longlong
_pi = ctypes.POINTER(ctypes.c_int)
_pd = ctypes.POINTER(ctypes.c_double)
_pll = ctypes.POINTER(ctypes.c_longlong)

# name -> (restype, argtypes)
SIGNATURES = {
    "wnfea_abi_version": (_i, []),
    "wnfea_device_info": (_i, [_pll]),
    "wnfea_hbm_stream_read": (_i, [_p, _ll, _p, _i, _p]),
    "wnfea_c3d10_ke": (_i, [_i, _p, _p, _p, _p, _p, _p, _p]),
    "wnfea_c3d10_ke_variant": (_i, [_i]),
    "wnfea_c3d10_fint": (_i, [_i, _i, _p, _p, _p, _p, _p, _i, _p, _p, _p, _p, _i, _p]),
    "wnfea_c3d10_stress": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "wnfea_adjoint_grad": (_i, [_i, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "wnfea_beam_forces": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _i, _p]),
    "wnfea_beam_tangent": (_i, [_i, _p, _p, _p, _p, _d, _p, _p]),
    "wnfea_beam_ke_linear": (_i, [_i, _p, _p, _p, _p, _p]),
    "wnfea_bsr_assemble_elem": (_i, [_p, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p]),
    "wnfea_rcm": (_i, [_i, _p, _p, _p, _p]),
    "wnfea_node_graph": (_ll, [_i, _i, _i, _p, _p, _p, _p, _ll]),
    "wnfea_contributions": (_ll, [_i, _i, _i, _p, _p, _p, _p, _p, _ll]),
    "wnfea_node_incidence": (_i, [_i, _i, _i, _p, _p, _p]),
    "wnfea_sa_aggregate": (_ll, [_i, _p, _p, _p]),
    "wnfea_sa_tentative": (_i, [_i, _i, _i, _p, _i, _p, _p, _p, _p, _p]),
    "wnfea_host_bsr_spgemm": (_ll, [_i, _i, _i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i]),
    "wnfea_host_bsr_transpose": (_i, [_i, _i, _i, _i, _p, _p, _p, _p, _p, _p]),
    "wnfea_host_bsr_symmetrize": (_ll, [_i, _i, _p, _p, _p, _i]),
    "wnfea_host_bsr_spmv": (_i, [_i, _i, _p, _p, _p, _p, _p, _i]),
    "wnfea_host_morton_order": (_i, [_i, _p, _p, _p]),
    "wnfea_host_mf_patch_sizes": (_ll, [_i, _i, _p, _p, _i, _pi]),
    "wnfea_host_mf_patches": (_ll, [_i, _i, _p, _p, _i, _i, _p, _p, _p, _p, _p, _p]),
    "wnfea_bsr_create": (_p, [_i, _i, _i, _i, _p, _p, _i, _i]),
    "wnfea_bsr_destroy": (None, [_p]),
    "wnfea_bsr_info": (_i, [_p, _pll]),
    "wnfea_bsr_tune": (_i, [_p, _i, _i, _i]),
    "wnfea_bsr_tune_precision": (_i, [_p, _i, _i, _i, _i]),
    "wnfea_bsr_set_values": (_i, [_p, _p, _i, _p]),
    "wnfea_bsr_get_values": (_i, [_p, _p, _p]),
    "wnfea_bsr_assemble": (_i, [_p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "wnfea_bsr_diag": (_i, [_p, _p, _p, _p]),
    "wnfea_bsr_equilibrate": (_i, [_p, _p, _p, _p, _i, _i, _p]),
    "wnfea_bsr_spmv": (_i, [_p, _i, _p, _p, _d, _d, _p]),
    "wnfea_bsr_spmv_dot": (_i, [_p, _i, _p, _p, _pd, _p]),
    "wnfea_amg_create": (_p, [_i, _i]),
    "wnfea_amg_destroy": (None, [_p]),
    "wnfea_amg_set_level": (_i, [_p, _i, _p, _i, _i, _p, _p, _p, _p, _d, _d, _i, _i, _i, _p, _p, _p, _p, _p, _p, _i]),
    "wnfea_amg_set_coarse": (_i, [_p, _i, _p]),
    "wnfea_amg_finalize": (_i, [_p, _i, _i, _i]),
    "wnfea_amg_tune_precision": (_i, [_p, _i, _i, _i, _i]),
    "wnfea_amg_vcycle": (_i, [_p, _i, _p, _p, _p]),
    "wnfea_amg_level_cheb": (_i, [_p, _i, _i, _i, _p]),
    "wnfea_amg_set_coarse_fp64": (_i, [_p, _i]),
    "wnfea_residual_noalloc": (_i, [_p, _p, _p, _p, _p, _p]),
    "wnfea_solver_create": (_p, []),
    "wnfea_solver_destroy": (None, [_p]),
    "wnfea_solver_set_graphs": (_i, [_p, _i]),
    "wnfea_pcg": (_i, [_p, _p, _i, _i, _p, _p, _i, _p, _p, _p, _d, _i, _i, _pi, _pd, _p]),
    "wnfea_ir": (_i, [_p, _p, _p, _i, _p, _p, _i, _p, _p, _d, _d, _i, _i, _i, _i, _pi, _pi, _pd, _p]),
    "wnfea_residual": (_i, [_p, _p, _p, _p, _p, _pd, _p]),
    "wnfea_mf_create": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, ctypes.POINTER(_p), _p]),
    "wnfea_mf_destroy": (None, [_p]),
    "wnfea_mf_apply": (_i, [_p, _p, _p, _p]),
    "wnfea_mf_residual": (_i, [_p, _p, _p, _p, _p, _pd, _p]),
    "wnfea_pcg_mf": (_i, [_p, _p, _i, _p, _p, _i, _p, _p, _p, _d, _i, _i, _pi, _pd, _p]),
    "wnfea_amg_attach_mf": (_i, [_p, _p, _p, _p]),
    "wnfea_mf_affine": (_i, [_p]),
    "wnfea_mf_set_patch": (_i, [_p, _i, _p]),
    "wnfea_mf_info": (_i, [_p, _pll]),
    "wnfea_pcg_multi": (_i, [_p, _p, _p, _i, _p, _i, _p, _p, _d, _i, _i, _pi, _pd, _p]),
    "wnfea_bsr_spmm": (_i, [_p, _i, _i, _p, _p, _p]),
}

_LIB = None


class HipError(RuntimeError):
    """Non-zero status from the HIP library."""


_STATUS = {-1: "bad argument", -2: "bad size", -3: "bad mode", -4: "not ready", -5: "bad mesh",
           -6: "no convergence"}


def load():
    """Load libwnfea_hip.so (after torch).  Raises if it was not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    import torch  # noqa: F401  (binds the process to torch's HIP runtime first)
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(the solve path has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(code: int, what: str) -> None:
    if code != 0:
        msg = _STATUS.get(code, f"HIP error {code}")
        raise HipError(f"{what} failed: {msg} ({code})")


def stream_of(device=None) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream


def dptr(t) -> int:
    """Device pointer of a contiguous CUDA(HIP) tensor."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return t.data_ptr()


def hptr(a: np.ndarray | None) -> int | None:
    """Host pointer of a C-contiguous numpy array."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("expected a C-contiguous array")
    return a.ctypes.data


def require_gpu():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the solve path runs on the GPU only; no device is visible")
    return torch.device("cuda", torch.cuda.current_device())
