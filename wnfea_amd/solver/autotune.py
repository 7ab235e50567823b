"""SpMV launch-shape autotuner with an on-disk cache.

The sliced-ELL SpMV has three launch knobs (csrc/sell.hpp): wavefronts per
slice (1, 2, 4), blocks in flight per lane (1, 2, 4) and the XCD-contiguous
slice mapping.  ``autotune_spmv`` times every combination on the actual
matrix with HIP events on the launch stream, keeps the fastest, applies it to
the matrix handle and caches the choice per (device, block shape, size,
precision).  The reference only has fixed launch heuristics
(``wnfea/solver/fast_kernels.py:243-260``).
"""
from __future__ import annotations

import itertools
import json
import os

_DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "wnfea_amd", "autotune.json")


def _cache_path() -> str:
    return os.environ.get("WNFEA_AUTOTUNE_CACHE", _DEFAULT_CACHE)


def _load_cache() -> dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(cache: dict) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only costs a re-tune next time


# bumped whenever the SpMV kernels or the timing method change, so that cached
# choices of an older build are not reused
_KERNEL_REV = 5

# GPU cycles of the spin kernel that holds the stream while the timed launches
# are queued (about 1 ms at the shader clock), and the fastest shader clock of
# the device (Hz): a hold of H cycles lasts at least H / _CLOCK_HZ seconds
_HOLD_CYCLES = 2_000_000
_CLOCK_HZ = 2.4e9


def _timed(fn, reps: int) -> float:
    """Average device seconds per call of ``fn`` (which launches on torch's
    current stream).  A spin kernel holds the stream while the host queues
    the ``reps`` calls, so that short kernels are timed back to back on the
    device instead of at the rate the host can launch them (a 13 us SpMV is
    otherwise as fast as a Python-level launch and the shapes time alike).
    The host's queueing time is measured: when it was not well inside the
    hold (the device may have drained the queue and waited for the host),
    the measurement is repeated with a hold sized from it."""
    import time
    import torch
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    hold = _HOLD_CYCLES
    for _ in range(4):
        torch.cuda.synchronize()
        torch.cuda._sleep(hold)
        a.record()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        t_queue = time.perf_counter() - t0
        b.record()
        torch.cuda.synchronize()
        if t_queue * _CLOCK_HZ <= 0.5 * hold:
            break
        hold = int(4 * t_queue * _CLOCK_HZ)
    return a.elapsed_time(b) / 1e3 / reps


def _key(K, precision: str) -> str:
    import torch
    name = torch.cuda.get_device_name(K.device)
    info = K.info()
    return (f"rev{_KERNEL_REV}|{name}|{K.br}x{K.bc}|rows={K.n_rows}|nnzb={K.nnzb}|lanes={info['lanes']}|"
            f"{precision}")


def time_config(K, precision: str, waves: int, unroll: int, xcd: bool, reps: int = 20) -> float:
    """Average seconds per SpMV launch for one launch shape."""
    import torch
    K.tune(waves, xcd, unroll, precision=precision)
    dt = torch.float32 if precision == "fp32" else torch.float64
    x = torch.ones(K.shape[1], dtype=dt, device=K.device)
    y = torch.empty(K.shape[0], dtype=dt, device=K.device)
    for _ in range(2):
        K.spmv(x, y, precision=precision)
    return _timed(lambda: K.spmv(x, y, precision=precision), reps)


def autotune_spmv(K, precision: str = "fp32", reps: int = 20, use_cache: bool = True) -> dict:
    """Pick and apply the fastest launch shape for K's `precision` values
    ("fp32" also sets the FP16 values' shape); returns the choice and timings."""
    key = _key(K, precision)
    cache = _load_cache() if use_cache else {}
    if key in cache:
        c = cache[key]
        K.tune(c["waves"], c["xcd_remap"], c["unroll"], precision=precision)
        return dict(c, cached=True)
    table = []
    for waves, unroll, xcd in itertools.product((1, 2, 4), (1, 2, 4), (True, False)):
        table.append({"waves": waves, "unroll": unroll, "xcd_remap": xcd,
                      "seconds": time_config(K, precision, waves, unroll, xcd, reps)})
    best = min(table, key=lambda r: r["seconds"])
    K.tune(best["waves"], best["xcd_remap"], best["unroll"], precision=precision)
    choice = {"waves": best["waves"], "unroll": best["unroll"], "xcd_remap": best["xcd_remap"],
              "seconds": best["seconds"]}
    if use_cache:
        cache[key] = choice
        _store_cache(cache)
    return dict(choice, cached=False, table=table)


def _cycle_key(amg, precision: str) -> str:
    import torch
    name = torch.cuda.get_device_name(amg.fine.device)
    shape = ",".join(f"{lv.A.shape[0]}/{lv.A.nnz}" for lv in amg.levels)
    mf = "mf" if getattr(amg, "matrix_free_smoothing", False) else "asm"
    return f"cycle-rev{_KERNEL_REV}|{name}|{shape}|{mf}|{precision}"


def autotune_cycle(amg, precision: str = "fp32", reps: int = 10, use_cache: bool = True) -> dict:
    """Pick and apply the SpMV launch shape of the V-cycle for one value
    precision by timing whole cycles (HIP events on the launch stream): with
    the finest level matrix-free the cycle's products are the coarse-level
    and transfer SpMVs, whose best shape differs from the fine matrix's
    (I-beam FP32: 4 wavefronts per slice, 1 block in flight, 3 % faster per
    solve than the fine matrix's 2 x 2).  Cached per hierarchy shape."""
    import torch
    key = _cycle_key(amg, precision)
    cache = _load_cache() if use_cache else {}
    if key in cache:
        c = cache[key]
        amg.set_tune(precision, c["waves"], c["unroll"], c["xcd_remap"])
        return dict(c, cached=True)
    dt = torch.float32 if precision != "fp64" else torch.float64
    n = amg.levels[0].A.shape[0]
    r = torch.randn(n, device=amg.fine.device).to(dt)
    z = torch.empty_like(r)
    table = []
    for waves, unroll in itertools.product((1, 2, 4), (1, 2)):
        amg.set_tune(precision, waves, unroll, True)
        for _ in range(2):
            amg.vcycle(r, z)
        table.append({"waves": waves, "unroll": unroll, "xcd_remap": True,
                      "seconds": _timed(lambda: amg.vcycle(r, z), reps)})
    best = min(table, key=lambda t: t["seconds"])
    amg.set_tune(precision, best["waves"], best["unroll"], True)
    choice = {k: best[k] for k in ("waves", "unroll", "xcd_remap", "seconds")}
    if use_cache:
        cache[key] = choice
        _store_cache(cache)
    return dict(choice, cached=False, table=table)
