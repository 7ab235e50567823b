"""Native-kernel entry points under the reference's module name
(``wnfea/solver/fast_kernels.py``), for callers that import them from there.

* ``compute_internal_forces_fast`` (``fast_kernels.py:359-448``): device beam
  and C3D10 forces (``beam_solver.compute_internal_forces_fast``).
* ``hip_spmv_bsr6x6`` (``fast_kernels.py:765-832``): block SpMV on the device
  (``bsr_matrix.hip_spmv_bsr6x6``).
* ``get_hip_device_summary`` (``fast_kernels.py:225-240``): the device
  summary from ``wnfea_device_info``, as a dict.
* ``get_optimal_launch_config`` (``fast_kernels.py:243-260``): kept for
  callers; this library picks its launch shapes itself (autotuned per
  matrix), so the pair is only advisory.
"""
from __future__ import annotations

from .. import _lib
from .beam_solver import compute_internal_forces_fast  # noqa: F401
from .bsr_matrix import hip_spmv_bsr6x6  # noqa: F401


def get_hip_device_summary() -> dict:
    """CUs, wavefront size, gfx arch number, L2 bytes, LDS bytes per block and
    HIP runtime version of the current device."""
    out = (_lib.ctypes.c_longlong * 6)()
    _lib.check(_lib.load().wnfea_device_info(out), "wnfea_device_info")
    keys = ("compute_units", "wavefront_size", "gfx_arch", "l2_bytes", "lds_bytes_per_block", "hip_runtime")
    return dict(zip(keys, (int(v) for v in out)))


def get_optimal_launch_config(n_items: int, item_type: str = "beam") -> tuple[int, int]:
    """(block_size, grid_size) for one item per thread: 256-thread blocks (four
    wavefronts) and enough blocks to cover every item."""
    block = 256
    return block, max(1, (int(n_items) + block - 1) // block)
