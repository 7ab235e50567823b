"""The reference's AMG interface (BlockBeamAMGPreconditioner, the 6-DOF
sparse assembly) on top of the device SA-AMG, checked the way the
reference's own tests check it (``test_gpu_amg.py:95-140``,
``test_mixed_precision.py:105-140, 360-380``) and against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def frame():
    from wnfea_amd.mesh.lattice import beam_props, generate_beam_lattice
    model, _ = generate_beam_lattice(4, lateral_sigma=500.0, compression_mean=2.0e3, compression_sigma=500.0)
    return model, beam_props(model)


def test_sparse_assembly_matches_oracle(frame):
    from oracle import beam_oracle as B
    from wnfea_amd.solver.beam_solver import assemble_global_system_sparse
    model, props = frame
    K, F = assemble_global_system_sparse(model)
    Ko = B.assemble(len(model.mesh_nodes), model.mesh_elements,
                    B.ke_linear(model.mesh_nodes, model.mesh_elements, props), model.fixed_dof_mask(6), sym=False)
    assert K.shape == Ko.shape and abs(K - Ko).max() <= 1e-12 * abs(Ko).max()
    fixed = model.fixed_dof_mask(6)
    Fo = model.load_vector(6)
    Fo[fixed] = 0.0
    np.testing.assert_array_equal(F, Fo)


def test_vcycles_fp64_fp32_and_tri_precision(frame):
    from wnfea_amd.solver.amg_preconditioner import BlockBeamAMGPreconditioner
    from wnfea_amd.solver.beam_solver import assemble_global_system_sparse
    model, _ = frame
    A, _ = assemble_global_system_sparse(model)
    amg64 = BlockBeamAMGPreconditioner(model, A=A, device="hip", working_dtype=np.float64)
    amg32 = BlockBeamAMGPreconditioner(model, A=A, device="hip", working_dtype=np.float32)
    assert amg64.gpu_handle is not None and len(amg64.levels) >= 2
    rng = np.random.default_rng(42)
    r, s = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    z64, z32 = amg64.matvec(r), amg32.matvec(r)
    assert z64.dtype == np.float64 and z32.dtype == np.float64
    assert float(r @ z64) > 0 and float(r @ z32) > 0
    assert abs(float(s @ z64) - float(r @ amg64.matvec(s))) <= 1e-10 * abs(float(r @ z64))
    assert _rel(z32, z64) < 1e-4
    for lv in amg32.levels:
        assert lv.A.dtype == np.float32 and lv.inv_diag.dtype == np.float32
        if lv.P is not None:
            assert lv.P.dtype == np.float32 and lv.R.dtype == np.float32
    amg32.enable_tri_precision(switch_tol=1e-2)
    z16 = amg32.apply_adaptive(r, current_res_rel=0.5)
    assert amg32.last_precision_used == "float16" and z16.dtype == np.float64
    assert _rel(z16, z64) < 5e-2
    amg32.apply_adaptive(r, current_res_rel=1e-4)
    assert amg32.last_precision_used == "float32"
    assert amg32.fp16_vcycle_count == 1 and amg32.fp32_vcycle_count >= 2


def test_solve_pcg_and_scipy_cg(frame):
    import scipy.sparse.linalg as sla
    from wnfea_amd.solver.amg_preconditioner import BlockBeamAMGPreconditioner
    from wnfea_amd.solver.beam_solver import assemble_global_system_sparse
    model, _ = frame
    A, F = assemble_global_system_sparse(model)
    u_ref = sla.spsolve(A.tocsc(), F)
    amg = BlockBeamAMGPreconditioner(model, A=A, working_dtype=np.float32)
    for mode in (0, 1, 2):
        u, it, rel = amg.solve_pcg(F, rtol=1e-10, max_iter=500, precision_mode=mode)
        assert rel <= 1e-9 and _rel(u, u_ref) <= 1e-7, mode
    its = []
    u2, info = sla.cg(A, F, rtol=1e-10, maxiter=500, M=amg, callback=lambda xk: its.append(1))
    assert info == 0 and _rel(u2, u_ref) <= 1e-7 and len(its) < 300


def test_solid_system_through_the_reference_interface(c0_model, c0_oracle):
    """A 3-DOF C3D10 operator given as A works too."""
    from wnfea_amd.solver.amg_preconditioner import BlockBeamAMGPreconditioner
    _, mr, _ = c0_model
    K, F, u_ref = c0_oracle
    amg = BlockBeamAMGPreconditioner(mr, A=K, working_dtype=np.float32, coarse_size=2000)
    u, it, rel = amg.solve_pcg(F, rtol=1e-9, max_iter=300)
    assert rel <= 1e-8 and _rel(u, u_ref) <= 1e-6 and it < 100
