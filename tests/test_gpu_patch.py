"""Patch layout of the matrix-free C3D10 operator on the device (elements in
Morton order, one workgroup per patch with its nodal values and element
forces in LDS, a node-ordered kernel adding each node's partial sums): its
products apply the oracle's assembled K for every patch size and equal the
per-element path to FP64 rounding; results are bitwise reproducible; the
V-cycle with patch-layout smoothing is the numpy mirror's cycle; a
hierarchy attached before a relayout falls back to the per-element path
until attached again; solves keep their iteration counts."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def _mesh(kind):
    from wnfea_amd.mesh.rcm import apply_rcm_to_model
    from wnfea_amd.mesh.solid_generator import generate_c3d10_ibeam, generate_c3d10_structured_block
    if kind == "ibeam":
        m, _ = generate_c3d10_ibeam(n_layers=12, length=0.12)
    else:
        m, _ = generate_c3d10_structured_block(1.0, 0.1, 0.1, 40, 4, 4, tip_load_total=-1000.0,
                                               tets_per_hex=5 if kind == "box5" else 6, load_direction="z")
    return apply_rcm_to_model(m)[0]


@pytest.mark.parametrize("kind", ["box5", "box6", "ibeam"])
def test_patch_products_apply_the_oracle_matrix(kind):
    import torch
    from oracle import oracle as O
    from wnfea_amd import _lib
    from wnfea_amd.solver.solid_solver import MatrixFreeOperator, assemble_c3d10_bsr, true_residual
    m = _mesh(kind)
    print(f"patch products on {kind}", flush=True)
    system = assemble_c3d10_bsr(m)
    op = MatrixFreeOperator(system)
    try:
        info = op.patch_info
        P0 = info["elems"]  # the default layout
        assert P0 == 64 and info["patches"] == -(-len(m.solid_elements) // P0)
        E, nu = m.element_moduli()
        K, _ = O.assemble_csr(m.mesh_nodes, m.solid_elements, E, nu, m.fixed_dof_mask(3), m.load_vector(3))
        rng = np.random.default_rng(21)
        u = rng.standard_normal(system.n_dofs)
        ud = torch.as_tensor(u, device="cuda")
        ref = K @ u
        y0 = op.apply(ud)
        assert _rel(y0.cpu().numpy(), ref) <= 1e-13
        assert torch.equal(op.apply(ud), y0)  # deterministic
        ys = {}
        for elems in (0, 128, 192, 256):
            try:
                op.set_patch(elems)
            except _lib.HipError as e:  # an FP64 patch of 192 or 256 elements may not fit the LDS
                assert elems >= 192 and "bad size" in str(e)
                continue
            assert op.patch_info["elems"] == elems
            ys[elems] = op.apply(ud).cpu().numpy()
            assert _rel(ys[elems], ref) <= 1e-13, elems
            assert _rel(ys[elems], y0.cpu().numpy()) <= 1e-14, elems
        op.set_patch(P0)
        assert torch.equal(op.apply(ud), y0)  # the rebuilt layout is the same layout
        # the fused residual (mode 2, with its dot partials)
        assert abs(op.residual(ud) - true_residual(system, ud)) <= 1e-12 * true_residual(system, ud)
    finally:
        op.close()


def test_patch_smoothing_cycles_and_relayout_fallback():
    import torch
    from vcycle_mirror import vcycle
    from wnfea_amd.solver.solid_solver import (MatrixFreeOperator, assemble_c3d10_bsr, build_amg, solve_mixed,
                                               true_residual)
    m = _mesh("ibeam")
    print("patch smoothing cycles on a 12-layer I-beam", flush=True)
    system = assemble_c3d10_bsr(m, want_fp64_equilibrated=True)
    amg = build_amg(system, storage=("fp64", "fp32"), autotune=False, theta=0.1, theta_decay=0.4, coarse_size=500)
    op = MatrixFreeOperator(system)
    P0 = op.patch_info["elems"]  # the default layout
    try:
        assert P0 > 0 and len(amg.levels) >= 3
        r = np.random.default_rng(22).standard_normal(system.n_dofs)
        rd = torch.as_tensor(r, device="cuda")
        z_ref = vcycle(amg.levels, amg.coarse_inv, r, deg=amg.level_degrees, ratio=amg.ratio)
        op.attach_smoother(amg)
        z64 = amg.vcycle(rd).cpu().numpy()
        z32 = amg.vcycle(rd.float()).double().cpu().numpy()
        assert _rel(z64, z_ref) <= 1e-12 and _rel(z32, z_ref) <= 1e-4
        it = {}
        for prec in ("fp64", "fp32"):
            # in the patch layout the cycle's last smoothing step also forms
            # the CG's rho; the per-element runs below use the rho pass
            u, it[prec], _, rel, _ = solve_mixed(system, amg, tol=1e-8, amg_precision=prec, operator=op)
            assert rel <= 1e-8 and true_residual(system, u) <= 1.5e-8, prec
        # per-element path (set_patch re-attaches the hierarchy)
        op.set_patch(0)
        z64e = amg.vcycle(rd).cpu().numpy()
        z32e = amg.vcycle(rd.float()).double().cpu().numpy()
        assert _rel(z64e, z64) <= 1e-13 and _rel(z32e, z32) <= 1e-5
        for prec in ("fp64", "fp32"):
            _, it_e, _, rel, _ = solve_mixed(system, amg, tol=1e-8, amg_precision=prec, operator=op)
            assert rel <= 1e-8 and abs(it_e - it[prec]) <= 1, (prec, it_e, it[prec])
        # a relayout behind the hierarchy's back: the FP32 cycle keeps its
        # per-element copies (the attachment was made without a patch layout)
        op.lib.wnfea_mf_set_patch(op.handle, P0, None)
        torch.cuda.synchronize()
        assert np.array_equal(amg.vcycle(rd.float()).double().cpu().numpy(), z32e)
        assert np.array_equal(amg.vcycle(rd).cpu().numpy(), z64)  # FP64 uses the operator's own layout
        amg.attach_matrix_free(op)
        assert np.array_equal(amg.vcycle(rd.float()).double().cpu().numpy(), z32)
    finally:
        op.close()
        amg.close()
