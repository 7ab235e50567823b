"""Benchmark of the C3D10 solve path on one MI355X.

Workload (default): the C3D10 I-beam, 3 m long, 0.2 x 0.02 flanges, 0.012 web,
0.3 deep, 300 hex layers split into 6 tets each (351 585 nodes, 1 054 755
DOF, 208 800 elements), root clamped, -50 kN shear + 5 kN m torsion at the
tip.  One step = one solve of the resident system from u = 0 to
||f - K u|| / ||f|| <= 1e-8 (FP64-verified) with the mixed-precision solver:
FP64 conjugate gradients on the matrix-free K preconditioned by a
smoothed-aggregation AMG V-cycle that runs in FP32 on the equilibrated
operator S K S (the finest level's smoothing steps matrix-free).

The printed JSON line carries value = seconds per solve (lower is better),
the same solve with the all-FP64 AMG-PCG and the speedup with a per-part
breakdown, the FP32 / FP64 3x3 BSR SpMV bandwidth (the metric's SpMV
figure), a roofline object for the dominant kernel of the timed solve (the
FP32 6x6 Chebyshev smoothing step of AMG level 1, three launches per
iteration; with the read rate a plain streaming kernel reaches on the same
device, the practical ceiling next to the datasheet peak), the displacement and von Mises errors against the CPU FP64
oracle, and a CPU baseline (the FP64 oracle on the host cores: a bounded
sample on the I-beam, extrapolated, and a measured solve to tolerance of the
19 683-DOF box).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "s to rel-res 1e-8 + BSR SpMV GB/s; mixed-prec vs FP64 PCG, same MI355X"
HBM_PEAK_GBS = 8000.0


def spmv_bytes(nnzb: int, nb: int, b: int, sv: int, sx: int) -> int:
    """Algorithmic bytes of one BSR(b) SpMV: values + block column indices +
    row pointers + x read once + y written once."""
    return nnzb * (b * b * sv + 4) + 4 * (nb + 1) + 2 * b * nb * sx


def cheb_bytes(nnzb: int, n: int, b: int, s: int) -> int:
    """Algorithmic bytes of one Chebyshev smoothing step on a BSR(b) level of
    n scalar rows with s-byte values and vectors: the matrix as in
    spmv_bytes, and per scalar row x, b, D^-1 and d read, d and x' written."""
    return nnzb * (b * b * s + 4) + 4 * (n // b + 1) + 6 * n * s


def time_level_cheb(amg, level: int, precision: str, reps: int) -> float:
    """Average device time (s) of one Chebyshev smoothing-step launch of an
    AMG level, HIP events on the stream the library launches on."""
    import torch
    from wnfea_amd import _lib
    L = _lib.load()
    vt = {"fp64": 0, "fp32": 1}[precision]
    stream = torch.cuda.current_stream()
    _lib.check(L.wnfea_amg_level_cheb(amg.handle, level, vt, 4, _lib.stream_of()), "wnfea_amg_level_cheb")
    st, en = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    st.record(stream)
    _lib.check(L.wnfea_amg_level_cheb(amg.handle, level, vt, reps, _lib.stream_of()), "wnfea_amg_level_cheb")
    en.record(stream)
    torch.cuda.synchronize()
    return st.elapsed_time(en) / 1e3 / reps


def stream_ceiling(reps: int = 10) -> dict:
    """Read rate of the library's HBM streaming-read probe (eight 16-byte
    non-temporal loads in flight per thread, 512-thread blocks) over a 1 GiB
    buffer, the best of three grid sizes, HIP events on the launch stream:
    the practical ceiling of a bandwidth-bound kernel on this device, next
    to the datasheet's 8 TB/s."""
    import torch
    from wnfea_amd import _lib
    L = _lib.load()
    nbytes = 1 << 30
    a = torch.zeros(nbytes // 4, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    per_grid = {}
    for grid in (2048, 4096, 8192):
        out = torch.empty(grid * 512, dtype=torch.float32, device="cuda")

        def run():
            _lib.check(L.wnfea_hbm_stream_read(_lib.dptr(a), nbytes, _lib.dptr(out), grid, _lib.stream_of()),
                       "wnfea_hbm_stream_read")
        run()
        run()
        st, en = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        st.record(stream)
        for _ in range(reps):
            run()
        en.record(stream)
        torch.cuda.synchronize()
        per_grid[grid] = round(nbytes / (st.elapsed_time(en) / 1e3 / reps) / 1e9, 1)
    del a
    torch.cuda.empty_cache()
    best = max(per_grid, key=per_grid.get)
    return {"gbs": per_grid[best], "grid": best, "per_grid_gbs": per_grid, "bytes": nbytes}


# Element stiffness kernel cost per C3D10 element, the dense convention of a
# B^T D B contraction: per Gauss point D B (6 x 6 x 30 multiply-adds) and
# B^T (D B) (30 x 30 x 6), 2 flop each, 4 points; bytes: 40 B connectivity,
# 10 nodes x 24 B coordinates, E and nu (16 B), and the 900-value FP64 Ke.
KE_FLOP_PER_ELEMENT = 4 * 2 * (6 * 6 * 30 + 30 * 30 * 6)
KE_BYTES_PER_ELEMENT = 40 + 240 + 16 + 900 * 8


def time_ke(system, reps: int) -> dict:
    """Device time of the element stiffness kernel on every element of the
    system (both kernel variants: vector ALU and matrix cores), HIP events on
    the launch stream; the process default (vector ALU) is restored."""
    import torch
    from wnfea_amd import _lib
    L = _lib.load()
    mesh = system.mesh
    M = mesh["n_elements"]
    Ke = torch.empty((M, 30, 30), dtype=torch.float64, device=system.K.device)
    bad = torch.zeros(1, dtype=torch.int32, device=system.K.device)
    stream = torch.cuda.current_stream()
    out = {}
    try:
        for variant, name in ((0, "valu"), (1, "mfma")):
            _lib.check(L.wnfea_c3d10_ke_variant(variant), "wnfea_c3d10_ke_variant")

            def run():
                _lib.check(L.wnfea_c3d10_ke(M, _lib.dptr(mesh["conn"]), _lib.dptr(mesh["nodes"]),
                                            _lib.dptr(mesh["E"]), _lib.dptr(mesh["nu"]), _lib.dptr(Ke),
                                            _lib.dptr(bad), _lib.stream_of()), "wnfea_c3d10_ke")
            run()
            st, en = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            st.record(stream)
            for _ in range(reps):
                run()
            en.record(stream)
            torch.cuda.synchronize()
            t = st.elapsed_time(en) / 1e3 / reps
            out[name] = {"us": round(t * 1e6, 1),
                         "tflops_fp64": round(M * KE_FLOP_PER_ELEMENT / t / 1e12, 2),
                         "frac_of_78.6_tflops": round(M * KE_FLOP_PER_ELEMENT / t / 78.6e12, 3),
                         "gbs": round(M * KE_BYTES_PER_ELEMENT / t / 1e9, 1),
                         "frac_of_hbm": round(M * KE_BYTES_PER_ELEMENT / t / 1e9 / HBM_PEAK_GBS, 3)}
    finally:
        L.wnfea_c3d10_ke_variant(0)
    out["elements"] = M
    out["flop_per_element"] = KE_FLOP_PER_ELEMENT
    out["bytes_per_element"] = KE_BYTES_PER_ELEMENT
    del Ke
    torch.cuda.empty_cache()
    return out


def jacobi_to_tol(system, tol: float, max_total: int) -> dict:
    """FP64 Jacobi-PCG on the device (the reference's algorithm), restarted on
    the true FP64 residual until ||f - K u|| <= tol ||f|| or max_total
    iterations; reports whether it converged."""
    import torch
    from wnfea_amd.solver.solid_solver import pcg_device, true_residual
    fn = float(system.F.norm())
    b, u, total, rel, runs = system.F, None, 0, float("inf"), 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while total < max_total and runs < 8:
        runs += 1
        rt = 0.5 * tol * fn / float(b.norm())
        x, it, _, _ = pcg_device(system, b=b, rtol=rt, max_iter=max_total - total, preconditioner="jacobi",
                                 precision="fp64", check_every=64)
        total += it
        u = x if u is None else u + x
        prev, rel = rel, true_residual(system, u)
        if rel <= tol or rel > 0.5 * prev:
            break
        b = system.F - system.K.spmv(u)
    torch.cuda.synchronize()
    return {"seconds": round(time.perf_counter() - t0, 4), "iterations": total, "cg_runs": runs, "rel_res": rel,
            "converged": rel <= tol}


CHEB_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "amg_level1_cheb_ibeam_1.05M_hbm_traffic.md")


def measured_traffic(path: str, precision: str):
    """HBM bytes per launch from a committed rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE summary (counters scaled by a 256 MiB copy kernel), or None
    when that summary is absent."""
    import re
    try:
        txt = open(path).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch\s+fp32=(\d+)\s+fp64=(\d+)", txt, re.M)
    if not m:
        return None
    return int(m.group(1) if precision == "fp32" else m.group(2))


def time_spmv(K, precision: str, reps: int) -> float:
    """Average device time (s) of one SpMV launch, HIP events on the launch stream."""
    import torch
    dt = torch.float32 if precision == "fp32" else torch.float64
    n = K.shape[0]
    x = torch.randn(n, dtype=torch.float64, device=K.device).to(dt)
    y = torch.zeros(n, dtype=dt, device=K.device)
    for _ in range(3):
        K.spmv(x, y, precision=precision)
    st, en = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    st.record()
    for _ in range(reps):
        K.spmv(x, y, precision=precision)
    en.record()
    torch.cuda.synchronize()
    return st.elapsed_time(en) / 1e3 / reps


def _oracle_pcg_to_tol(O, K, F, tol: float, max_total: int):
    """Oracle FP64 Jacobi-PCG restarted on its true residual (CPU SpMV) until
    ||F - K x|| <= tol ||F||; (x, iterations, true relative residual)."""
    fn = np.linalg.norm(F)
    b, x, total, rel = F, np.zeros_like(F), 0, float("inf")
    for _ in range(8):
        d, it, _ = O.pcg(K, b, rtol=0.5 * tol * fn / np.linalg.norm(b), max_iter=max_total - total)
        total += it
        x = x + d
        b = F - O.csr_spmv(K, x)
        prev, rel = rel, np.linalg.norm(b) / fn
        if rel <= tol or total >= max_total or rel > 0.5 * prev:
            break
    return x, total, rel


def oracle_check(model, U, mesh=None, owner=None, n_groups: int = 0) -> dict:
    """The CPU FP64 oracle's check of device solutions of one C3D10 model:
    U holds one displacement vector (3 DOF per node, the model's node order)
    or one per column; the relative residual of each on the oracle's own K and
    CPU SpMV (with ``owner``, the node -> mesh map of a merged model, per
    mesh), and the device cell / nodal von Mises fields of the first column
    against the oracle's stress field of the same displacements."""
    import torch
    from oracle import oracle as O
    from wnfea_amd.solver.solid_solver import compute_c3d10_stress_field
    E, nu = model.element_moduli()
    K, F = O.assemble_csr(model.mesh_nodes, model.solid_elements, E, nu, model.fixed_dof_mask(3),
                          model.load_vector(3))
    Uh = U.detach().cpu().numpy() if isinstance(U, torch.Tensor) else np.asarray(U)
    cols = Uh.reshape(len(F), -1)
    out = {}
    if owner is not None:
        r = F - O.csr_spmv(K, cols[:, 0])
        own3 = np.repeat(np.asarray(owner), 3)
        r2 = np.bincount(own3, weights=r * r, minlength=n_groups)
        f2 = np.bincount(own3, weights=F * F, minlength=n_groups)
        rel = np.sqrt(r2) / np.where(f2 > 0, np.sqrt(f2), 1.0)
        out["rel_res_on_oracle_K_max_over_meshes"] = float(rel.max())
    out["cols"] = cols.shape[1]
    _, cv, nv = compute_c3d10_stress_field(model, torch.as_tensor(cols[:, 0].copy()), mesh)
    _, rcv, rnv = O.stress_field(model.mesh_nodes, model.solid_elements, E, nu, cols[:, 0])
    out["von_mises_kernel_vs_oracle_same_u_max_rel"] = float(max(np.abs(nv - rnv).max() / np.abs(rnv).max(),
                                                                  np.abs(cv - rcv).max() / np.abs(rcv).max()))
    return out, K, F


def cpu_baseline(model, u_mixed, u_ref, vm_device, gpu_jacobi_iters: int | None, sample_iters: int = 60):
    """The CPU leg: the FP64 oracle (OpenMP C restatement of the reference's
    path) on the host cores.

    Baseline: assembly and a bounded sample of `sample_iters` Jacobi-PCG
    iterations on the same I-beam system, extrapolated to the iterations the
    same algorithm takes to 1e-8 (the device Jacobi-PCG, restarted on its
    true residual); and a measured oracle solve to 1e-8 of the 19 683-DOF
    box.  Check: the oracle's K gives the residual of the mixed solution and
    of the FP64 reference solution; the displacement error of the mixed
    solution against the oracle-verified FP64 solution; the device von Mises
    fields against the oracle's on the FP64 solution (error) and on the same
    mixed displacements (kernel parity)."""
    from oracle import oracle as O
    from wnfea_amd.mesh.rcm import apply_rcm_to_model
    from wnfea_amd.mesh.solid_generator import generate_c3d10_structured_block
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    E, nu = model.element_moduli()
    t0 = time.perf_counter()
    K, F = O.assemble_csr(model.mesh_nodes, model.solid_elements, E, nu, model.fixed_dof_mask(3),
                          model.load_vector(3))
    t_asm = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, it, rel = O.pcg(K, F, rtol=1e-30, max_iter=sample_iters)
    t_it = (time.perf_counter() - t0) / max(it, 1)
    need = gpu_jacobi_iters
    out = {"unit": "s", "cores": threads, "kind": "port",
           "sample": (f"oracle FP64 Jacobi-PCG (the reference's algorithm, solid_solver.py:197-308) on the same "
                      f"I-beam system: {it} iterations sampled ({t_it * 1e3:.1f} ms/iteration), assembly "
                      f"{t_asm:.1f} s; value EXTRAPOLATED from those {it} iterations to the {need} iterations the "
                      f"same algorithm takes to a true residual of 1e-8 (device Jacobi-PCG restarted on its FP64 "
                      f"residual)"),
           "extrapolated": True}
    out["value"] = round(t_it * need, 3) if need else None
    out["seconds_per_iteration"] = t_it
    out["sampled_iterations"] = it
    out["assembly_sec"] = t_asm
    # a measured CPU solve to tolerance: the 19 683-DOF cantilever box
    box, _ = generate_c3d10_structured_block(1.0, 0.1, 0.1, 40, 4, 4, tip_load_total=-1000.0, tets_per_hex=6,
                                             load_direction="z")
    box, _, _ = apply_rcm_to_model(box)
    Eb, nub = box.element_moduli()
    t0 = time.perf_counter()
    Kb, Fb = O.assemble_csr(box.mesh_nodes, box.solid_elements, Eb, nub, box.fixed_dof_mask(3), box.load_vector(3))
    t_asm_b = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, itb, relb = _oracle_pcg_to_tol(O, Kb, Fb, 1e-8, 200000)
    out["box_19683_dof_measured"] = {"assembly_sec": round(t_asm_b, 3), "solve_sec": round(time.perf_counter() - t0, 3),
                                     "iterations": itb, "rel_res": relb, "solver": "oracle FP64 Jacobi-PCG, restarted "
                                     "on the true residual"}
    # the check against the oracle
    uh, urh = u_mixed.cpu().numpy(), u_ref.cpu().numpy()
    fn = np.linalg.norm(F)
    _, cv_ref, nv_ref = O.stress_field(model.mesh_nodes, model.solid_elements, E, nu, urh)
    _, cv_same, nv_same = O.stress_field(model.mesh_nodes, model.solid_elements, E, nu, uh)
    cv_d, nv_d = vm_device
    err = {"reference_solution": "device FP64 AMG-PCG to the FP64 residual floor, verified on the oracle's K",
           "rel_res_mixed_on_oracle_K": float(np.linalg.norm(F - O.csr_spmv(K, uh)) / fn),
           "rel_res_reference_on_oracle_K": float(np.linalg.norm(F - O.csr_spmv(K, urh)) / fn),
           "displacement_rel_err": float(np.linalg.norm(uh - urh) / np.linalg.norm(urh)),
           "von_mises_nodal_rel_err": float(np.linalg.norm(nv_d - nv_ref) / np.linalg.norm(nv_ref)),
           "von_mises_cell_rel_err": float(np.linalg.norm(cv_d - cv_ref) / np.linalg.norm(cv_ref)),
           "von_mises_kernel_vs_oracle_same_u_max_rel": float(max(np.abs(nv_d - nv_same).max() / np.abs(nv_same).max(),
                                                               np.abs(cv_d - cv_same).max() / np.abs(cv_same).max()))}
    return out, err


def spmv_only(system, info, reps: int, log) -> int:
    """SpMV launches alone: `reps` FP32 then `reps` FP64 SpMVs on the resident
    I-beam matrix, then 5 copies of a 256 MiB FP32 buffer (known HBM bytes:
    256 MiB read + 256 MiB written per copy) to calibrate the HBM counters."""
    import torch
    t32 = time_spmv(system.K, "fp32", reps)
    t64 = time_spmv(system.K, "fp64", reps)
    n = 64 * 2 ** 20
    a = torch.ones(n, dtype=torch.float32, device=system.K.device)
    c = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        c.copy_(a)
    torch.cuda.synchronize()
    nnzb, nb = info["nnzb"], system.n_nodes
    print(json.dumps({"spmv_fp32_us": t32 * 1e6, "spmv_fp64_us": t64 * 1e6,
                      "bytes_fp32": spmv_bytes(nnzb, nb, 3, 4, 4), "bytes_fp64": spmv_bytes(nnzb, nb, 3, 8, 8),
                      "copy_bytes_each_way": 4 * n}), flush=True)
    log("[bench] spmv-only done")
    return 0


def mf_bytes(n_elements: int, n_nodes: int, s: int) -> dict:
    """Algorithmic bytes of the matrix-free kernels on an affine C3D10 mesh
    with s-byte values: the element kernel reads the connectivity (40 B), the
    quadrature data (10 values), E and nu (16 B) per element and every node's
    value and scaling once (6 values), and writes 30 values per element; the
    Chebyshev-step gather reads the incidence codes (40 B) and the 30 element
    values per element, and per node its incidence pointer, Dirichlet byte
    and five operands (b, D^-1, x, d, s), and writes two (x', d')."""
    elem = n_elements * (40 + 10 * s + 16 + 30 * s) + n_nodes * 6 * s
    gather = n_elements * (40 + 30 * s) + n_nodes * (4 + 1 + 15 * s + 6 * s)
    return {"elem": elem, "gather_cheb": gather}


def mf_patch_bytes(n_elements: int, n_nodes: int, n_local: int, n_patch: int, stride: int, s: int,
                   folded: bool) -> dict:
    """Algorithmic bytes of the patch layout's two kernels with s-byte values.
    The patch kernel reads per element its 16-bit local connectivity (20 B),
    quadrature record (folded FP32: J^-1 and the weighted moduli, 12 values;
    FP64: J^-1 and det J / 24, 10 values, plus the element id and E, nu,
    20 B) and slot codes (20 B), per patch its stride-padded local-node ids
    and list starts (6 B each), every node's value and scaling once (6
    values), and writes one partial sum (3 values) per local node.  The node
    kernel reads per node its 16-byte head and Dirichlet byte, per local node
    its partial sum (3 values), and the epilogue's operands: a Chebyshev step
    reads b, D^-1, x, d, s (15 values) and writes x', d' (6); the FP64
    product reads u (3) and writes y (3)."""
    per_elem = 20 + 20 + (12 * s if folded else 10 * s + 20)
    patch = n_elements * per_elem + n_patch * stride * 6 + n_local * 3 * s + n_nodes * 6 * s
    node = n_nodes * (16 + 1) + n_local * 3 * s
    return {"patch": patch, "node_cheb": node + n_nodes * 21 * s, "node_defect": node + n_nodes * 12 * s,
            "node_apply": node + n_nodes * 6 * s}


def mf_only(system, model, reps: int, log, patch: int = -1) -> int:
    """Matrix-free kernels alone on the resident I-beam: `reps` FP64 operator
    applications (element kernel + gather), then `reps` FP32 V-cycles with the
    finest level's smoothing and defect matrix-free and `reps` with them on
    the assembled matrix, then 5 copies of a 256 MiB FP32 buffer to calibrate
    the HBM counters."""
    import torch
    from wnfea_amd.solver.solid_solver import MatrixFreeOperator, build_amg
    op = MatrixFreeOperator(system)
    if patch >= 0:
        op.set_patch(patch)
    amg = build_amg(system, storage=("fp64", "fp32"))
    u = torch.randn(system.n_dofs, dtype=torch.float64, device=system.K.device)
    r = torch.randn(system.n_dofs, dtype=torch.float32, device=system.K.device)

    def timed(fn):
        for _ in range(3):
            fn()
        st, en = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        st.record()
        for _ in range(reps):
            fn()
        en.record()
        torch.cuda.synchronize()
        return st.elapsed_time(en) / 1e3 / reps

    y = torch.empty_like(u)
    t_apply = timed(lambda: op.apply(u, y))
    z = torch.empty_like(r)
    t_asm = timed(lambda: amg.vcycle(r, z))
    op.attach_smoother(amg)
    t_mf = timed(lambda: amg.vcycle(r, z))
    n = 64 * 2 ** 20
    a = torch.ones(n, dtype=torch.float32, device=system.K.device)
    c = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        c.copy_(a)
    torch.cuda.synchronize()
    ne, nn = system.mesh["n_elements"], system.n_nodes
    pinfo = op.patch_info
    pb = {p: mf_patch_bytes(ne, nn, pinfo["local_nodes"], pinfo["patches"], pinfo["stride"], s, folded=(s == 4))
          for p, s in (("fp32", 4), ("fp64", 8))}
    print(json.dumps({"mf64_apply_us": t_apply * 1e6, "vcycle32_matrix_free_us": t_mf * 1e6,
                      "vcycle32_assembled_us": t_asm * 1e6, "affine": op.affine, "layout": pinfo,
                      "bytes_fp32": mf_bytes(ne, nn, 4), "bytes_fp64": mf_bytes(ne, nn, 8),
                      "patch_bytes_fp32": pb["fp32"], "patch_bytes_fp64": pb["fp64"],
                      "copy_bytes_each_way": 4 * n}), flush=True)
    op.close()
    amg.close()
    log("[bench] mf-only done")
    return 0


def run_batch(args, log) -> int:
    """64 seeded uneven cantilever boxes (nx in [10,60], ny = nz in [3,8],
    2k-60k DOF each) solved as one block-diagonal system to 1e-8 per mesh,
    against the same boxes solved one after another on the same GPU."""
    import torch
    from wnfea_amd.mesh.rcm import apply_rcm_to_model
    from wnfea_amd.solver.batch import BatchedSolids, generate_box_batch
    from wnfea_amd.solver.solid_solver import assemble_c3d10_bsr, build_amg, solve_mixed
    t0 = time.perf_counter()
    batch = generate_box_batch(64, seed=1)
    models = [b[0] for b in batch]
    n_dofs = sum(b[1]["n_dofs"] for b in batch)
    log(f"[bench] batch: 64 boxes, {n_dofs} DOF total ({time.perf_counter() - t0:.1f}s)")
    kw = {"ratio": args.amg_ratio, **omega_kw(args)}
    if args.amg_theta is not None:
        kw["theta"] = args.amg_theta
    if args.amg_theta_decay is not None:
        kw["theta_decay"] = args.amg_theta_decay
    bs = BatchedSolids(models, amg_kw=kw)
    log(f"[bench] batched setup {bs.timings}")
    for _ in range(args.warmup):
        bs.solve(1e-8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        u, iters, per, _ = bs.solve(1e-8)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / args.steps
    log(f"[bench] batched solve {t_batch * 1e3:.1f} ms, {iters} iterations, max per-mesh rel {per.max():.2e}")
    chk, _, _ = oracle_check(bs.merged, u, bs.system.mesh, owner=bs.owner_host, n_groups=bs.n_meshes)
    chk.pop("cols")
    log(f"[bench] oracle check {chk}")
    bs.close()
    # the same meshes one by one (each its own system, hierarchy and matrix-free
    # operator, as in the batch); solve time only
    from wnfea_amd.solver.solid_solver import MatrixFreeOperator
    systems = []
    for m in models:
        mr, _, _ = apply_rcm_to_model(m)
        s_ = assemble_c3d10_bsr(mr)
        a_ = build_amg(s_, storage=("fp32",), ratio=args.amg_ratio)
        op_ = MatrixFreeOperator(s_)
        op_.attach_smoother(a_)
        systems.append((s_, a_, op_))
    for s_, a_, op_ in systems:
        solve_mixed(s_, a_, tol=1e-8, operator=op_)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seq_iters = []
    for s_, a_, op_ in systems:
        seq_iters.append(solve_mixed(s_, a_, tol=1e-8, operator=op_)[1])
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    for _, a_, op_ in systems:
        a_.close()
        op_.close()
    log(f"[bench] one-by-one solves {t_seq * 1e3:.1f} ms (iterations {min(seq_iters)}..{max(seq_iters)})")
    print(json.dumps({
        "metric": "s to rel-res 1e-8 per mesh, 64 uneven C3D10 boxes in one system",
        "value": round(t_batch, 5), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_batch * 1e3, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 CG / f32 AMG", "data": "synthetic (seeded box sizes, deterministic loads)",
        "config": {"workload": "c3d10_64_boxes_batch", "n_dofs_total": n_dofs, "seed": 1,
                   "parallelism": "single-gpu"},
        "dof_per_s": round(n_dofs / t_batch, 1), "iterations": iters, "max_per_mesh_rel_res": float(per.max()),
        "one_by_one_seconds": round(t_seq, 5), "speedup_batched_vs_one_by_one": round(t_seq / t_batch, 2),
        "err_vs_oracle": chk, "setup": bs.timings}), flush=True)
    return 0


def run_adjoint(args, log) -> int:
    """I-beam with per-element moduli E_e = 210 GPa (0.2 + 0.8 U(0,1)) (seed 2):
    compliance f.u and mean tip z-displacement, their gradients with respect to
    every E_e (one forward and one adjoint solve), the CPU FP64 reference
    gradient (oracle element matrices with the same u and lambda) and central
    differences (h = 1e-4 E_e) on 20 seeded elements (seed 3)."""
    import torch
    from oracle import oracle as O
    from wnfea_amd.mesh.rcm import apply_rcm_to_model
    from wnfea_amd.mesh.solid_generator import generate_c3d10_ibeam
    from wnfea_amd.solver.adjoint import C3D10Adjoint, tip_vector
    model, meta = generate_c3d10_ibeam(n_layers=args.layers, length=3.0 * args.layers / 300)
    model, n2o, o2n = apply_rcm_to_model(model)
    M = len(model.solid_elements)
    E = 210e9 * (0.2 + 0.8 * np.random.default_rng(2).uniform(size=M))
    model.solid_E, model.solid_nu = E, np.full(M, 0.3)
    tip = o2n[meta["tip_nodes"]]
    c = tip_vector(model, tip)
    t0 = time.perf_counter()
    # the library's strength threshold: with per-element moduli spread over
    # a factor of five the I-beam's thresholds (0.1, decay 0.5) are slower
    # here (27.7 / 52.9 ms against 26.8 / 51.0 ms)
    prob = C3D10Adjoint(model, tol=1e-10, amg_precision="fp32", amg_kw=omega_kw(args))
    prob.set_moduli(E)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    prob.compliance_and_grad(E)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    C, gC, u = prob.compliance_and_grad(E)
    torch.cuda.synchronize()
    t_c = time.perf_counter() - t0
    t0 = time.perf_counter()
    T, gT, _, lam = prob.response_and_grad(E, c)
    torch.cuda.synchronize()
    t_t = time.perf_counter() - t0
    log(f"[bench] compliance {C:.6e} + gradient {t_c * 1e3:.1f} ms; tip {T:.6e} + gradient {t_t * 1e3:.1f} ms")
    uh, lh = u.cpu().numpy(), lam.cpu().numpy()
    chk, Ko, Fo = oracle_check(model, u, prob.system.mesh)
    chk.pop("cols")
    chk["rel_res_forward_on_oracle_K"] = float(np.linalg.norm(Fo - O.csr_spmv(Ko, uh)) / np.linalg.norm(Fo))
    chk["rel_res_adjoint_on_oracle_K"] = float(np.linalg.norm(c - O.csr_spmv(Ko, lh)) / np.linalg.norm(c))
    del Ko
    log(f"[bench] oracle check {chk}")
    t0 = time.perf_counter()
    refC = O.adjoint_grad(model.mesh_nodes, model.solid_elements, E, model.solid_nu, uh, uh)
    refT = O.adjoint_grad(model.mesh_nodes, model.solid_elements, E, model.solid_nu, uh, lh)
    t_cpu = time.perf_counter() - t0
    gCh, gTh = gC.cpu().numpy(), gT.cpu().numpy()
    errC = float(np.linalg.norm(gCh - refC) / np.linalg.norm(refC))
    errT = float(np.linalg.norm(gTh - refT) / np.linalg.norm(refT))
    rng = np.random.default_rng(3)
    fd_err = []
    t0 = time.perf_counter()
    for e in rng.choice(M, 20, replace=False):
        # central difference of C from forward solves of K(E +- h e_e) (FP64 floor)
        fd = prob.compliance_central_difference(E, int(e), 1e-4 * E[e])
        fd_err.append((abs(fd - gCh[e]) / abs(gCh[e]), abs(fd - gCh[e]) / np.abs(gCh).max(), abs(gCh[e])))
    t_fd = time.perf_counter() - t0
    prob.close()
    fd_rel = [a for a, _, _ in fd_err]
    fd_scaled = [b for _, b, _ in fd_err]
    log(f"[bench] gradient vs CPU FP64: compliance {errC:.2e}, tip {errT:.2e}; FD error / max|g| "
        f"{max(fd_scaled):.2e}, per-element relative median {np.median(fd_rel):.2e} max {max(fd_rel):.2e}")
    print(json.dumps({
        "metric": "compliance + tip-displacement adjoint gradients, C3D10 I-beam, per-element E",
        "value": round(t_c + t_t, 5), "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round((t_c + t_t) * 1e3, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 CG / f32 AMG, f64 gradients",
        "data": "synthetic (seeded moduli, deterministic loads)",
        "config": {"workload": "c3d10_ibeam_adjoint_per_element_E", "n_dofs": meta["n_dofs"], "n_elements": M,
                   "seed_E": 2, "seed_fd": 3, "parallelism": "single-gpu"},
        "compliance_seconds": round(t_c, 5), "tip_seconds": round(t_t, 5), "setup_seconds": round(t_setup, 3),
        "grad_rel_err_vs_cpu_fp64": {"compliance": errC, "tip": errT},
        "fd_20_elements": {"max_err_over_max_abs_grad": float(max(fd_scaled)),
                           "median_rel_err": float(np.median(fd_rel)), "max_rel_err": float(max(fd_rel)),
                           "h": "1e-4 E_e", "seconds": round(t_fd, 2),
                           "method": "C(E+h)-C(E-h) = f.(d+ - d-) with u+- = u0 + d+-, each d solved on the "
                                     "device-assembled perturbed system to the FP64 floor"},
        "cpu_reference_gradient_seconds": round(t_cpu, 3), "err_vs_oracle": chk,
        "solves": prob.solves}), flush=True)
    return 0


# SA strength threshold on the I-beam: 0.1 on the finest level times 0.5 per
# coarser level (sweep with the prolongator smoothing factor 1.6: 40
# iterations; 0.3 / 0.4 / 0.45 / 0.55 / 0.6 take 44 / 41 / 41 / 40 / 39, the
# last one with a coarsest level of 11 694 rows, close to the 12 000 rows
# above which a fourth level appears and the iterations rise to 67)
IBEAM_THETA_DECAY = 0.5


def omega_kw(args) -> dict:
    """The SA prolongator smoothing factor when --amg-omega is given (else
    the solver's default for the problem type)."""
    return {} if args.amg_omega is None else {"omega": args.amg_omega}


def tune_spmv(K, shape: str | None, log) -> dict:
    """SpMV launch shapes for the FP32 and the FP64 values of K: fixed by
    `shape` ("waves,unroll,xcd_remap" for both, or "fp32 shape:fp64 shape"),
    or each autotuned."""
    from wnfea_amd.solver.autotune import autotune_spmv
    if shape:
        parts = shape.split(":")
        out = {}
        for prec, text in zip(("fp32", "fp64"), (parts[0], parts[-1])):
            w, u, x = (int(v) for v in text.split(","))
            K.tune(w, bool(x), u, precision=prec)
            out[prec] = {"waves": w, "unroll": u, "xcd_remap": bool(x)}
        return out
    t0 = time.perf_counter()
    out = {}
    for prec in ("fp64", "fp32"):
        c = autotune_spmv(K, prec, use_cache=False)
        out[prec] = {k: c[k] for k in ("waves", "unroll", "xcd_remap", "seconds")}
    log(f"[bench] autotune {out} ({time.perf_counter() - t0:.1f}s)")
    return out


def apply_tune(amg, tunes: dict) -> None:
    """The cycle's launch shapes from the fine matrix's (FP32/FP16 and FP64 values)."""
    for p in ("fp32", "fp64"):
        t = tunes[p]
        amg.set_tune(p, t["waves"], t["unroll"], t["xcd_remap"])


def run_multirhs(args, log) -> int:
    """C2a: 32 seeded tip load cases (forces N(0, 20 kN), moments N(0, 2 kN m)
    per component, seed 0) on the I-beam, solved in groups of 8 columns (one
    SpMM per iteration per group) against 32 single-column mixed solves (the
    fastest single-column form: matrix-free CG operator and finest smoother)."""
    import torch
    from wnfea_amd.mesh.rcm import apply_rcm_to_model
    from wnfea_amd.mesh.solid_generator import generate_c3d10_ibeam, ibeam_tip_load_cases
    from wnfea_amd.solver.solid_solver import assemble_c3d10_bsr, build_amg, solve_mixed, solve_multi
    model, meta = generate_c3d10_ibeam(n_layers=args.layers, length=3.0 * args.layers / 300)
    model, n2o, o2n = apply_rcm_to_model(model)
    system = assemble_c3d10_bsr(model)
    tunes = tune_spmv(system.K, args.spmv_shape, log)
    tune = tunes["fp32"]
    # the I-beam's strength thresholds (as the single-load-case benchmark)
    amg = build_amg(system, storage=("fp32",), waves=tune["waves"], unroll=tune["unroll"],
                    xcd_remap=tune["xcd_remap"],
                    theta=args.amg_theta if args.amg_theta is not None else 0.1,
                    theta_decay=args.amg_theta_decay if args.amg_theta_decay is not None else IBEAM_THETA_DECAY,
                    **omega_kw(args))
    apply_tune(amg, tunes)
    B = ibeam_tip_load_cases(model.mesh_nodes, o2n[meta["tip_nodes"]], n_cases=32, seed=0)
    B[system.fixed_mask] = 0.0
    Bt = torch.as_tensor(B, device="cuda")
    solve_multi(system, amg, Bt[:, :8], tol=1e-8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    U, iters, rels, _ = solve_multi(system, amg, Bt, tol=1e-8)
    torch.cuda.synchronize()
    t_multi = time.perf_counter() - t0
    log(f"[bench] 32 load cases in groups of 8: {t_multi * 1e3:.1f} ms, iterations {iters.min()}..{iters.max()}, "
        f"max rel {rels.max():.2e}")
    # the single-column solves in their fastest form: CG and the finest
    # smoothing steps on the matrix-free operator
    from wnfea_amd.solver.solid_solver import MatrixFreeOperator
    op = MatrixFreeOperator(system)
    op.attach_smoother(amg)
    F0 = system.F
    t_single = 0.0
    err = 0.0
    try:
        for k in range(32):
            system.F = Bt[:, k].contiguous()
            if k == 0:
                solve_mixed(system, amg, tol=1e-8, operator=op)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            u, _, _, _, _ = solve_mixed(system, amg, tol=1e-8, operator=op)
            torch.cuda.synchronize()
            t_single += time.perf_counter() - t0
            err = max(err, float(torch.linalg.norm(u - U[:, k]) / torch.linalg.norm(u)))
    finally:
        system.F = F0
    amg.close()
    op.close()
    log(f"[bench] 32 single-column solves (matrix-free operators): {t_single * 1e3:.1f} ms; max rel diff {err:.1e}")
    chk, Ko, _ = oracle_check(model, U[:, :1], system.mesh)
    from oracle import oracle as O
    Uh = U.cpu().numpy()
    rel_cols = [float(np.linalg.norm(B[:, k] - O.csr_spmv(Ko, Uh[:, k])) / np.linalg.norm(B[:, k])) for k in range(32)]
    chk.pop("cols")
    chk["rel_res_on_oracle_K_max_over_cases"] = max(rel_cols)
    kmax = int(np.argmax(rel_cols))
    # the FP64 rounding floor of each case's residual, eps * || |K| |u_k| || / ||b_k||
    absK = abs(Ko)
    floor = [float(2.2e-16 * np.linalg.norm(absK @ np.abs(Uh[:, k])) / np.linalg.norm(B[:, k])) for k in range(32)]
    chk["worst_case"] = {"case": kmax, "rel_res_device": float(rels[kmax]), "rel_res_oracle": rel_cols[kmax],
                         "fp64_floor_estimate": floor[kmax]}
    chk["fp64_floor_estimate_max_over_cases"] = max(floor)
    log(f"[bench] oracle check {chk}")
    print(json.dumps({
        "metric": "s to rel-res 1e-8 for 32 load cases on the C3D10 I-beam (multi-RHS)",
        "value": round(t_multi, 5), "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(t_multi * 1e3, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 CG / f32 AMG", "data": "synthetic (seeded tip loads)",
        "config": {"workload": "c3d10_ibeam_32_load_cases", "n_dofs": meta["n_dofs"], "n_rhs": 32, "group": 8,
                   "seed": 0, "parallelism": "single-gpu"},
        "seconds_per_load_case": round(t_multi / 32, 6), "single_column_seconds_total": round(t_single, 5),
        "speedup_vs_single_column": round(t_single / t_multi, 2), "iterations_min_max": [int(iters.min()),
                                                                                        int(iters.max())],
        "max_rel_res": float(rels.max()), "max_rel_diff_vs_single": err, "err_vs_oracle": chk}), flush=True)
    return 0


def run_beams(args, log) -> int:
    """C4: cubic lattice of co-rotational beams (cells^3, 1 m members, r = 0.02 m,
    E = 70 GPa, nu = 0.33), base clamped, seeded lateral + compressive top
    loads (seed 4), Newton over the load steps with FP64 CG + FP32 AMG inner
    solves.  Also: the internal-force kernel alone (HIP events) and the numpy
    restatement of the reference's vectorised force path on the host."""
    import torch
    from oracle import beam_oracle as B
    from wnfea_amd.mesh.lattice import beam_props, generate_beam_lattice
    from wnfea_amd.mesh.rcm import apply_rcm_to_model
    from wnfea_amd.solver.beam_solver import BeamSystem
    t0 = time.perf_counter()
    model, meta = generate_beam_lattice(args.cells)
    model, _, _ = apply_rcm_to_model(model)
    bs = BeamSystem(model)
    U0 = torch.zeros(6 * bs.n_nodes, dtype=torch.float64, device="cuda")
    bs.assemble_tangent(U0)
    amg_kw = {}
    if args.amg_theta is not None:
        amg_kw["theta"] = args.amg_theta
    if args.amg_theta_decay is not None:
        amg_kw["theta_decay"] = args.amg_theta_decay
    amg_kw.update(omega_kw(args))
    bs.build_amg(**amg_kw)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    log(f"[bench] lattice {meta['n_nodes']} nodes {meta['n_beams']} beams; setup {t_setup:.1f}s "
        f"(AMG levels {[l['n'] for l in bs.amg.stats()['levels']]}, cycle launch shape "
        f"{ {k: v for k, v in (getattr(bs, 'cycle_tune', None) or {}).items() if k != 'table'} })")
    U, st = bs.solve_newton(n_load_steps=args.load_steps)
    t_newton = st["solve_sec"]
    log(f"[bench] newton {t_newton:.2f}s its/step {st['newton_iters']} pcg {st['pcg_iters']} "
        f"final |R| {st['final_residual']:.3e}")
    # force kernel alone
    reps = 20
    bs.internal_forces(U)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        bs.internal_forces(U)
    e1.record()
    torch.cuda.synchronize()
    t_f = e0.elapsed_time(e1) / 1e3 / reps
    M = bs.n_beams
    bytes_f = M * 296 + bs.n_nodes * (6 * 8 + 4) + 2 * M * 4
    props = beam_props(model)
    Uh = U.cpu().numpy()
    t0 = time.perf_counter()
    Fo = B.corot_forces(model.mesh_nodes, model.mesh_elements, props, Uh)
    t_cpu = time.perf_counter() - t0
    Fg = bs.internal_forces(U).cpu().numpy()
    err = float(np.linalg.norm(Fg - Fo) / np.linalg.norm(Fo))
    # equilibrium of the final state with the oracle's forces (full load)
    Fext = model.load_vector(6)
    free = ~model.fixed_dof_mask(6)
    eq_res = float(np.linalg.norm((Fext - Fo)[free]) / np.linalg.norm(Fext[free]))
    log(f"[bench] final state: ||F_ext - f_int_oracle(U)|| / ||F_ext|| on free DOFs {eq_res:.2e}")
    bs.close()
    log(f"[bench] force kernel {t_f * 1e6:.1f} us ({M / t_f / 1e6:.0f} M elem/s), numpy host {t_cpu * 1e3:.1f} ms, "
        f"rel diff {err:.1e}")
    print(json.dumps({
        "metric": "s for 20-step co-rotational Newton solve, 40^3 beam lattice (413k DOF)",
        "value": round(t_newton, 4), "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(t_newton * 1e3, 2), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 residual/Newton, f64 CG / f32 AMG", "data": "synthetic (seeded loads)",
        "config": {"workload": f"corot_beam_lattice_{args.cells}^3", "n_dofs": meta["n_dofs"], "n_beams": M,
                   "load_steps": args.load_steps, "seed": 4, "amg_theta": amg_kw.get("theta", 0.02),
                   "amg_theta_decay": amg_kw.get("theta_decay", 1.0), "chebyshev_ratio": 20.0,
                   "parallelism": "single-gpu"},
        "newton_iters_per_step": st["newton_iters"], "pcg_iters": st["pcg_iters"], "tangents": st["tangents"],
        "final_residual": st["final_residual"], "setup_sec": round(t_setup, 2),
        "err_vs_oracle": {"equilibrium_rel_res_with_oracle_forces": eq_res,
                          "force_kernel_rel_diff_vs_oracle": err},
        "cycle_tune": {k: v for k, v in (getattr(bs, "cycle_tune", None) or {}).items() if k != "table"},
        "force_kernel": {"seconds": t_f, "elements_per_s": M / t_f, "gbs": bytes_f / t_f / 1e9,
                         "bytes_per_call": bytes_f, "rel_diff_vs_oracle": err},
        "cpu_baseline": {"value": round(t_cpu, 4), "unit": "s per internal-force evaluation", "cores": 1,
                         "kind": "port", "sample": "numpy restatement of fast_kernels.py:451-537 on the same lattice"},
    }), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", choices=["ibeam", "box", "batch", "adjoint", "beams", "multirhs"], default="ibeam",
                    help="ibeam/box: the solve metric; batch: 64 uneven boxes in one system; "
                         "adjoint: compliance and tip-displacement gradients on the I-beam; "
                         "beams: nonlinear co-rotational 40^3-cell lattice, 20 load steps")
    ap.add_argument("--cells", type=int, default=40, help="lattice cells per side (beams)")
    ap.add_argument("--load-steps", type=int, default=20, help="load steps (beams)")
    ap.add_argument("--layers", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-jacobi", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--amg-precision", choices=["fp16", "fp32", "fp16_fine", "fp16_coarse", "fp16_smooth"],
                    default="fp32")
    ap.add_argument("--graphs", action="store_true",
                    help="replay PCG iteration chunks as HIP graphs (default: host launches, measured faster here)")
    ap.add_argument("--operator", choices=["matrix_free", "assembled"], default="matrix_free",
                    help="CG's operator K: matrix-free (quadrature data) or the assembled FP64 BSR matrix")
    ap.add_argument("--amg-theta", type=float, default=None,
                    help="SA strength threshold on the finest level (I-beam default 0.1, box 0.08, beam lattice 0.02)")
    ap.add_argument("--amg-ratio", type=float, default=30.0,
                    help="Chebyshev smoother interval [lmax/ratio, lmax] of the I-beam, box and batch hierarchies")
    ap.add_argument("--amg-omega", type=float, default=None,
                    help="SA prolongator smoothing factor (default: 1.6 for the solids, 4/3 for the beam lattice)")
    ap.add_argument("--amg-theta-decay", type=float, default=None,
                    help="factor on the threshold per coarser level (I-beam default 0.5, box 1.0)")
    ap.add_argument("--smoother-operator", choices=["matrix_free", "assembled"], default="matrix_free",
                    help="operator of the finest level's smoothing steps and defect in the V-cycle")
    ap.add_argument("--spmv-shape", default=None,
                    help="waves,unroll,xcd_remap of the SpMV instead of autotuning, for both value precisions "
                         "(e.g. 4,1,0) or as fp32:fp64 (e.g. 2,2,1:4,1,1)")
    ap.add_argument("--mf-only", action="store_true",
                    help="time the matrix-free operator and V-cycle kernels alone (for counter runs)")
    ap.add_argument("--spmv-only", action="store_true",
                    help="only the SpMV launches (and a 256 MiB copy for counter calibration), for rocprofv3 --pmc")
    ap.add_argument("--ke-only", action="store_true",
                    help="only the element stiffness kernels (both variants) on the I-beam, for rocprofv3")
    ap.add_argument("--cheb-only", action="store_true",
                    help="only the level-1 Chebyshev smoothing steps (FP32 then FP64) and a 256 MiB copy, "
                         "for rocprofv3 --pmc")
    ap.add_argument("--no-ir", action="store_true", help="skip the restart-IR (FP32 Krylov) scheme")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--cycle-shape", default=None,
                    help="fix the V-cycle's launch shapes (waves,unroll,xcd_remap for both, or 'fp32:fp64') "
                         "instead of timing them on whole cycles")
    ap.add_argument("--mf-patch", type=int, default=-1,
                    help="elements per patch of the matrix-free operator's patch layout (0: per-element "
                         "contributions and node gather; default: the library's layout)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the solve path has no multi-GPU implementation (it would shard "
              "by contiguous RCM row ranges with one halo exchange per SpMV); rerun with --gpus 1", file=sys.stderr)
        return 2
    import torch
    from wnfea_amd.mesh.rcm import apply_rcm_to_model
    from wnfea_amd.mesh.solid_generator import generate_c3d10_ibeam, generate_c3d10_structured_block
    from wnfea_amd.solver.autotune import autotune_spmv
    from wnfea_amd.solver.solid_solver import (SOLID_SA_OMEGA, assemble_c3d10_bsr, build_amg, ir_device, pcg_device,
                                               set_graphs, solve_mixed, true_residual)

    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    torch.cuda.set_device(0)
    set_graphs(args.graphs)
    if args.config == "batch":
        return run_batch(args, log)
    if args.config == "adjoint":
        return run_adjoint(args, log)
    if args.config == "beams":
        return run_beams(args, log)
    if args.config == "multirhs":
        return run_multirhs(args, log)
    t0 = time.perf_counter()
    if args.config == "ibeam":
        model, meta = generate_c3d10_ibeam(n_layers=args.layers, length=3.0 * args.layers / 300)
        workload = "c3d10_ibeam_3m_1.05M_dof" if args.layers == 300 else f"c3d10_ibeam_{args.layers}_layers"
    else:
        model, meta = generate_c3d10_structured_block(1.0, 0.1, 0.1, 40, 4, 4, tip_load_total=-1000.0,
                                                      tets_per_hex=6, load_direction="z")
        workload = "c3d10_box_40x4x4_19683_dof"
    t_mesh = time.perf_counter() - t0
    log(f"[bench] mesh {meta['n_dofs']} DOF {meta['n_elements']} elements ({t_mesh:.1f}s)")
    t0 = time.perf_counter()
    model, n2o, _ = apply_rcm_to_model(model)
    t_rcm = time.perf_counter() - t0
    t0 = time.perf_counter()
    system = assemble_c3d10_bsr(model, want_fp64_equilibrated=True)
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    info = system.K.info()
    log(f"[bench] rcm {t_rcm:.1f}s assembly {t_asm:.1f}s nnzb {info['nnzb']} slots {info['slots']}")
    ke = None
    if args.config == "ibeam" and not (args.spmv_only or args.mf_only or args.cheb_only):
        ke = time_ke(system, 5)
        log(f"[bench] element stiffness kernel {ke}")
    if args.ke_only:
        print(json.dumps(time_ke(system, 10)), flush=True)
        return 0
    if args.no_autotune and not args.spmv_shape:
        args.spmv_shape = "1,1,1"
    tunes = tune_spmv(system.K, args.spmv_shape, log)
    tune = tunes["fp32"]
    if args.spmv_only:
        return spmv_only(system, info, args.spmv_reps, log)
    if args.mf_only:
        return mf_only(system, model, args.spmv_reps, log, args.mf_patch)
    t0 = time.perf_counter()
    # strength thresholds per workload (sweep on the I-beam: 0.1 on the finest
    # level and 0.04 below it take 44 instead of 50 iterations)
    th = args.amg_theta if args.amg_theta is not None else (0.1 if args.config == "ibeam" else 0.08)
    thd = args.amg_theta_decay if args.amg_theta_decay is not None else (IBEAM_THETA_DECAY if args.config == "ibeam"
                                                                         else 1.0)
    amg = build_amg(system, waves=tune["waves"], unroll=tune["unroll"], xcd_remap=tune["xcd_remap"],
                    theta=th, theta_decay=thd, ratio=args.amg_ratio, **omega_kw(args))
    apply_tune(amg, tunes)
    t_amg = time.perf_counter() - t0
    st = amg.stats()
    log(f"[bench] amg setup {t_amg:.1f}s levels {[l['n'] for l in st['levels']]} "
        f"op-complexity {st['operator_complexity']:.2f}")

    # CG's operator: matrix-free K (default) or the assembled FP64 BSR K;
    # every result is verified with the assembled FP64 K
    from wnfea_amd.solver.solid_solver import MatrixFreeOperator
    mf_op = MatrixFreeOperator(system)
    if args.mf_patch >= 0:
        mf_op.set_patch(args.mf_patch)
    log(f"[bench] matrix-free layout {mf_op.patch_info}")
    main_op = mf_op if args.operator == "matrix_free" else None
    alt_op = None if args.operator == "matrix_free" else mf_op
    alt_name = "assembled" if args.operator == "matrix_free" else "matrix_free"
    cycle_tune = None
    if args.smoother_operator == "matrix_free":
        mf_op.attach_smoother(amg)
        # the cycle's products are now the coarse-level and transfer SpMVs:
        # their own launch shapes (timed on whole cycles)
        from wnfea_amd.solver.autotune import autotune_cycle
        cycle_tune = {}
        parts = args.cycle_shape.split(":") if args.cycle_shape else None
        for p, text in zip(("fp32", "fp64"), (parts[0], parts[-1]) if parts else (None, None)):
            if text:  # fixed, e.g. so that a profiled run launches the default run's kernels
                w, u, x = (int(v) for v in text.split(","))
                amg.set_tune(p, w, u, bool(x))
                cycle_tune[p] = {"waves": w, "unroll": u, "xcd_remap": bool(x), "fixed": True}
            else:
                c = autotune_cycle(amg, p)
                cycle_tune[p] = {k: c[k] for k in ("waves", "unroll", "xcd_remap", "seconds")}
        log(f"[bench] cycle launch shapes {cycle_tune}")

    if args.cheb_only:
        solve_mixed(system, amg, tol=1e-8, amg_precision="fp32", operator=main_op)  # fills the level vectors
        solve_mixed(system, amg, tol=1e-8, amg_precision="fp64", operator=main_op)
        lv1 = amg.stats()["levels"][1]
        nnzb1 = lv1["nnz"] // (lv1["block"] ** 2)
        t32 = time_level_cheb(amg, 1, "fp32", args.spmv_reps)
        t64 = time_level_cheb(amg, 1, "fp64", args.spmv_reps)
        n = 64 * 2 ** 20
        a = torch.ones(n, dtype=torch.float32, device=system.K.device)
        c = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(5):
            c.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"cheb_fp32_us": t32 * 1e6, "cheb_fp64_us": t64 * 1e6, "n": lv1["n"], "nnzb": nnzb1,
                          "bytes_fp32": cheb_bytes(nnzb1, lv1["n"], lv1["block"], 4),
                          "bytes_fp64": cheb_bytes(nnzb1, lv1["n"], lv1["block"], 8),
                          "copy_bytes_each_way": 4 * n, "cycle_tune": cycle_tune}), flush=True)
        return 0

    def mixed_solve():
        return solve_mixed(system, amg, tol=1e-8, amg_precision=args.amg_precision, operator=main_op)

    def fp64_solve():
        return solve_mixed(system, amg, tol=1e-8, amg_precision="fp64", operator=main_op)

    for _ in range(args.warmup):
        mixed_solve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = [mixed_solve() for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    u, iters, steps, rel, _ = res[-1]
    rel_true = true_residual(system, u)
    log(f"[bench] mixed (FP64 CG on {args.operator} K + {args.amg_precision} V-cycle, {args.smoother_operator} "
        f"finest smoother): {t_mixed * 1e3:.1f} ms/solve "
        f"iterations {iters} restarts {steps - 1} rel {rel_true:.2e}")
    # all-FP64 AMG-PCG on the same hierarchy
    fp64_solve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 2)):
        x64, it64, _, _, _ = fp64_solve()
    torch.cuda.synchronize()
    t_fp64 = (time.perf_counter() - t0) / max(1, args.steps // 2)
    rel64 = true_residual(system, x64)
    log(f"[bench] fp64 AMG-PCG ({args.operator} K): {t_fp64 * 1e3:.1f} ms/solve iters {it64} rel {rel64:.2e}")
    err_u = float(torch.linalg.norm(u - x64) / torch.linalg.norm(x64))
    others = {}
    for other in ("fp32", "fp16", "fp16_fine", "fp16_coarse", "fp16_smooth"):
        if other == args.amg_precision:
            continue
        solve_mixed(system, amg, tol=1e-8, amg_precision=other, operator=main_op)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        uo, ito, _, relo, _ = solve_mixed(system, amg, tol=1e-8, amg_precision=other, operator=main_op)
        torch.cuda.synchronize()
        t_other = time.perf_counter() - t0
        others[f"mixed_{other}_vcycle"] = {"seconds": round(t_other, 5), "iterations": ito, "rel_res": relo}
        log(f"[bench] mixed with {other} V-cycle: {t_other * 1e3:.1f} ms iters {ito} rel {relo:.2e}")
    # the mixed solve with its coarsest solve on the FP64 inverse (FP64 products,
    # FP32 vectors), the precision the reference reserves for the coarse solve
    amg.set_coarse_precision("fp64")
    solve_mixed(system, amg, tol=1e-8, amg_precision="fp32", operator=main_op)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 2)):
        uc, itc, _, relc, _ = solve_mixed(system, amg, tol=1e-8, amg_precision="fp32", operator=main_op)
    torch.cuda.synchronize()
    t_c64 = (time.perf_counter() - t0) / max(1, args.steps // 2)
    amg.set_coarse_precision("fp32")
    others["mixed_fp32_vcycle_fp64_coarse_solve"] = {"seconds": round(t_c64, 5), "iterations": itc, "rel_res": relc}
    log(f"[bench] mixed with the FP64 coarsest solve: {t_c64 * 1e3:.1f} ms iters {itc} rel {relc:.2e}")
    # the same two solves with the other operator for CG
    alt = {"operator": alt_name}
    for name, prec in (("mixed", args.amg_precision), ("fp64", "fp64")):
        solve_mixed(system, amg, tol=1e-8, amg_precision=prec, operator=alt_op)
        torch.cuda.synchronize()
        n_rep = max(1, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(n_rep):
            um, itm, _, _, _ = solve_mixed(system, amg, tol=1e-8, amg_precision=prec, operator=alt_op)
        torch.cuda.synchronize()
        tm = (time.perf_counter() - t0) / n_rep
        relm = true_residual(system, um)
        alt[name] = {"seconds": round(tm, 5), "iterations": itm, "rel_res": relm}
        log(f"[bench] {name} with the {alt_name} CG operator: {tm * 1e3:.1f} ms iters {itm} rel {relm:.2e}")
    alt["speedup_mixed_vs_fp64"] = round(alt["fp64"]["seconds"] / alt["mixed"]["seconds"], 3)
    # the same two solves with the other operator in the finest-level smoother
    alt_sm_name = "assembled" if args.smoother_operator == "matrix_free" else "matrix_free"
    amg.attach_matrix_free(mf_op if alt_sm_name == "matrix_free" else None)
    if alt_sm_name == "assembled":
        apply_tune(amg, tunes)  # the fine matrix's shapes for a cycle that uses it
    alt_sm = {"operator": alt_sm_name}
    for name, fn in (("mixed", mixed_solve), ("fp64", fp64_solve)):
        fn()
        torch.cuda.synchronize()
        n_rep = max(1, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(n_rep):
            um, itm, _, _, _ = fn()
        torch.cuda.synchronize()
        tm = (time.perf_counter() - t0) / n_rep
        relm = true_residual(system, um)
        alt_sm[name] = {"seconds": round(tm, 5), "iterations": itm, "rel_res": relm}
        log(f"[bench] {name} with the {alt_sm_name} smoother operator: {tm * 1e3:.1f} ms iters {itm} "
            f"rel {relm:.2e}")
    alt_sm["speedup_mixed_vs_fp64"] = round(alt_sm["fp64"]["seconds"] / alt_sm["mixed"]["seconds"], 3)
    amg.attach_matrix_free(mf_op if args.smoother_operator == "matrix_free" else None)
    if cycle_tune:
        for p_, c in cycle_tune.items():
            amg.set_tune(p_, c["waves"], c["unroll"], c["xcd_remap"])
    # the same two solves with the other launch mode (graph replay or host launches)
    set_graphs(not args.graphs)
    eager = {}
    for name, fn in (("mixed", mixed_solve), ("fp64", fp64_solve)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        eager[name] = round(time.perf_counter() - t0, 5)
    set_graphs(args.graphs)
    log(f"[bench] {'host launches' if args.graphs else 'graphs'}: {eager}")
    ir = None
    if not args.no_ir:
        # the reference's mixed scheme (solid_solver.py:271-300): FP32 Krylov on S K S inside FP64 refinement
        _, ir_outer, ir_inner, ir_rel, ir_t = ir_device(system, tol=1e-8, amg=amg, max_outer=30)
        ir = {"seconds": round(ir_t, 4), "outer": ir_outer, "inner": ir_inner, "rel_res": ir_rel,
              "converged": ir_rel <= 1e-8}
        log(f"[bench] restart IR (FP32 Krylov): {ir}")
    # the reference's algorithm on the GPU: FP64 Jacobi-PCG, restarted on its true residual
    jac = None
    if not args.no_jacobi:
        jac = jacobi_to_tol(system, 1e-8, 200000)
        log(f"[bench] fp64 Jacobi-PCG: {jac}")
    # per-iteration breakdown of the mixed and the all-FP64 solve: CG's product,
    # one V-cycle in each precision (device time), the rest (vector kernels and
    # reductions) from the measured solve times
    r64 = torch.randn(system.n_dofs, dtype=torch.float64, device=system.K.device)
    r32 = r64.float()

    def ev_time(fn, reps):
        for _ in range(3):
            fn()
        st_, en_ = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        st_.record()
        for _ in range(reps):
            fn()
        en_.record()
        torch.cuda.synchronize()
        return st_.elapsed_time(en_) / 1e3 / reps
    z32, z64, y64 = torch.empty_like(r32), torch.empty_like(r64), torch.empty_like(r64)
    t_vc32 = ev_time(lambda: amg.vcycle(r32, z32, storage=args.amg_precision if args.amg_precision == "fp32"
                                        else "fp32"), 20)
    t_vc64 = ev_time(lambda: amg.vcycle(r64, z64, storage="fp64"), 20)
    t_kp = ev_time(lambda: mf_op.apply(r64, y64), 20)
    rest_mixed = t_mixed / iters - t_kp - t_vc32
    rest_fp64 = t_fp64 / it64 - t_kp - t_vc64
    breakdown = {"cg_product_fp64_us": round(t_kp * 1e6, 1), "vcycle_fp32_us": round(t_vc32 * 1e6, 1),
                 "vcycle_fp64_us": round(t_vc64 * 1e6, 1),
                 "vector_kernels_and_host_mixed_us": round(rest_mixed * 1e6, 1),
                 "vector_kernels_and_host_fp64_us": round(rest_fp64 * 1e6, 1),
                 "iteration_mixed_us": round(t_mixed / iters * 1e6, 1),
                 "iteration_fp64_us": round(t_fp64 / it64 * 1e6, 1),
                 "vcycle_speedup_fp32_vs_fp64": round(t_vc64 / t_vc32, 3),
                 "speedup_bound_if_vcycle_exactly_2x": round((t_kp + rest_fp64 + t_vc64) /
                                                             (t_kp + rest_fp64 + 0.5 * t_vc64), 3),
                 "note": "FP64 CG keeps its product and vector work in FP64 in both solvers; only the V-cycle "
                         "changes precision"}
    log(f"[bench] breakdown {breakdown}")
    # the dominant kernel of the timed solve: the FP32 6x6 Chebyshev step of AMG
    # level 1 (a hierarchy with a smoothed level between the finest and the
    # dense coarsest one; otherwise the fine K's FP32 SpMV below)
    has_l1 = len(st["levels"]) >= 3
    if has_l1:
        lv1 = st["levels"][1]
        b1 = lv1["block"]
        nnzb1 = lv1["nnz"] // (b1 * b1)
        t_ch32 = time_level_cheb(amg, 1, "fp32", args.spmv_reps)
        t_ch64 = time_level_cheb(amg, 1, "fp64", args.spmv_reps)
        ch32, ch64 = cheb_bytes(nnzb1, lv1["n"], b1, 4), cheb_bytes(nnzb1, lv1["n"], b1, 8)
        log(f"[bench] level-1 Chebyshev step fp32 {t_ch32 * 1e6:.1f} us {ch32 / t_ch32 / 1e9:.0f} GB/s | "
            f"fp64 {t_ch64 * 1e6:.1f} us {ch64 / t_ch64 / 1e9:.0f} GB/s")
    # the practical streaming-read ceiling of this device (plain read probe)
    ceiling = stream_ceiling()
    log(f"[bench] HBM streaming-read probe {ceiling['gbs']} GB/s (grid {ceiling['grid']}; {ceiling['per_grid_gbs']})")
    # SpMV bandwidth of the FP32 and FP64 3x3 kernel (the metric's SpMV figure)
    nnzb, nb = info["nnzb"], system.n_nodes
    t32 = time_spmv(system.K, "fp32", args.spmv_reps)
    t64 = time_spmv(system.K, "fp64", args.spmv_reps)
    b32 = spmv_bytes(nnzb, nb, 3, 4, 4)
    b64 = spmv_bytes(nnzb, nb, 3, 8, 8)
    gbs32, gbs64 = b32 / t32 / 1e9, b64 / t64 / 1e9
    log(f"[bench] spmv fp32 {t32 * 1e6:.1f} us {gbs32:.0f} GB/s | fp64 {t64 * 1e6:.1f} us {gbs64:.0f} GB/s")
    # the matrix-free FP64 operator (CG's product): element kernel + gather
    y_mf = torch.empty_like(u)
    for _ in range(3):
        mf_op.apply(u, y_mf)
    st_ev, en_ev = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    st_ev.record()
    for _ in range(args.spmv_reps):
        mf_op.apply(u, y_mf)
    en_ev.record()
    torch.cuda.synchronize()
    t_mf = st_ev.elapsed_time(en_ev) / 1e3 / args.spmv_reps
    ne_mf = system.mesh["n_elements"]
    pinfo = mf_op.patch_info
    if pinfo["elems"]:
        pb = mf_patch_bytes(ne_mf, nb, pinfo["local_nodes"], pinfo["patches"], pinfo["stride"], 8, folded=False)
        b_mf = pb["patch"] + pb["node_apply"]
    else:
        b_mf = mf_bytes(ne_mf, nb, 8)["elem"] + ne_mf * (40 + 30 * 8) + nb * (4 + 1 + 3 * 8)
    mf_info = {"apply_fp64_us": round(t_mf * 1e6, 2), "algorithmic_bytes": b_mf,
               "gbs": round(b_mf / t_mf / 1e9, 1), "affine": mf_op.affine, "layout": pinfo}
    log(f"[bench] matrix-free FP64 apply {t_mf * 1e6:.1f} us {b_mf / t_mf / 1e9:.0f} GB/s algorithmic")
    # FP64 reference solution (device FP64 AMG-PCG to the FP64 residual floor)
    # and the device von Mises fields of the mixed solution, for the oracle check
    from wnfea_amd.solver.solid_solver import compute_c3d10_stress_field
    u_ref, it_ref, _, rel_ref, _ = solve_mixed(system, amg, tol=1e-12, amg_precision="fp64", operator=main_op,
                                               max_restarts=6)
    _, cv_d, nv_d = compute_c3d10_stress_field(model, u, system.mesh)
    cpu, err_oracle = None, None
    if not args.no_cpu:
        log("[bench] cpu baseline and oracle check ...")
        cpu, err_oracle = cpu_baseline(model, u, u_ref, (cv_d, nv_d),
                                       jac["iterations"] if jac and jac["converged"] else None)
        log(f"[bench] cpu baseline {cpu['value']} s ({cpu['cores']} threads); oracle check {err_oracle}")
    out = {
        "metric": METRIC,
        "value": round(t_mixed, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"f64 CG / {args.amg_precision}-stored f32 AMG",
        "data": "synthetic (structured C3D10 mesh generated in-process, deterministic loads)",
        "config": {"workload": workload, "n_dofs": meta["n_dofs"], "n_elements": meta["n_elements"],
                   "nnzb": nnzb, "tol": 1e-8, "solver": "FP64 CG + low-precision SA-AMG V-cycle",
                   "cg_operator": args.operator, "smoother_operator": args.smoother_operator,
                   "amg_theta": th, "amg_theta_decay": thd,
                   "amg_omega": args.amg_omega if args.amg_omega is not None else SOLID_SA_OMEGA,
                   "parallelism": "single-gpu"},
        "dof_per_s": round(meta["n_dofs"] / t_mixed, 1),
        "rel_res_fp64_verified": rel_true,
        "iterations": iters,
        "cg_restarts": steps - 1,
        "vcycle_storage": args.amg_precision,
        "fp64_amg_pcg": {"seconds": round(t_fp64, 5), "iterations": it64, "rel_res": rel64},
        "cg_operator": args.operator,
        "other_cg_operator": alt,
        "smoother_operator": args.smoother_operator,
        "other_smoother_operator": alt_sm,
        "speedup_mixed_vs_fp64_amg_pcg": round(t_fp64 / t_mixed, 3),
        "speedup_breakdown": breakdown,
        "err_vs_oracle": err_oracle,
        "fp64_jacobi_pcg": jac,
        "restart_ir_fp32_krylov": ir,
        **others,
        ("host_launch_seconds" if args.graphs else "graph_replay_seconds"): eager,
        "graphs": args.graphs,
        "speedup_mixed_vs_fp64_jacobi_pcg": round(jac["seconds"] / t_mixed, 2) if jac and jac["converged"] else None,
        "displacement_rel_err_mixed_vs_fp64": err_u,
        "spmv_gbs": {"fp32": round(gbs32, 1), "fp64": round(gbs64, 1), "fp32_us": round(t32 * 1e6, 2),
                     "fp64_us": round(t64 * 1e6, 2), "bytes_fp32": b32, "bytes_fp64": b64},
        "roofline": ({"bound": "hbm", "achieved": round(ch32 / t_ch32 / 1e9, 1), "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": round(ch32 / t_ch32 / 1e9 / HBM_PEAK_GBS, 4),
                      "traffic": measured_traffic(CHEB_TRAFFIC_FILE, "fp32") if workload.startswith("c3d10_ibeam")
                      else None,
                      "traffic_source": "committed rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE summary "
                                        "(profiles/amg_level1_cheb_ibeam_1.05M_hbm_traffic.md), not measured in this run",
                      "algorithmic_bytes": ch32, "avg_launch_us": round(t_ch32 * 1e6, 2),
                      "stream_ceiling_gbs": ceiling["gbs"],
                      "frac_of_stream_ceiling": round(ch32 / t_ch32 / 1e9 / ceiling["gbs"], 4),
                      "launches_per_iteration": 3,
                      "kernel": f"sell_spmv_kernel<6,6,float,float,"
                                f"{cycle_tune['fp32']['waves'] if cycle_tune else tune['waves']},"
                                f"{cycle_tune['fp32']['unroll'] if cycle_tune else tune['unroll']},OpCheb<6,float>> "
                                f"(FP32 Chebyshev smoothing step of AMG level 1: {lv1['n']} rows, {nnzb1} 6x6 blocks)",
                      "fp64_twin": {"avg_launch_us": round(t_ch64 * 1e6, 2), "algorithmic_bytes": ch64,
                                    "frac": round(ch64 / t_ch64 / 1e9 / HBM_PEAK_GBS, 4)}} if has_l1 else
                     {"bound": "hbm", "achieved": round(gbs32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": round(gbs32 / HBM_PEAK_GBS, 4), "traffic": None, "algorithmic_bytes": b32,
                      "stream_ceiling_gbs": ceiling["gbs"],
                      "frac_of_stream_ceiling": round(gbs32 / ceiling["gbs"], 4),
                      "avg_launch_us": round(t32 * 1e6, 2),
                      "kernel": "FP32 3x3 SpMV of the fine K (the hierarchy has no smoothed level between the "
                                "finest and the dense coarsest one)"}),
        "spmv_tune": tunes,
        "cycle_tune": cycle_tune,
        "matrix_free_operator": mf_info,
        "element_stiffness_kernel": ke,
        "setup": {"mesh_sec": round(t_mesh, 2), "rcm_sec": round(t_rcm, 2), "assembly_sec": round(t_asm, 2),
                  "amg_setup_sec": round(t_amg, 2), "amg": st},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
