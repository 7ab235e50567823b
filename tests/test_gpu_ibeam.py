"""The 1.05M-DOF C3D10 I-beam (3 m, 300 layers) at full size: properties
that do not need an oracle solve of this size.  The matrix-free operator
applies the assembled FP64 K and is symmetric; the mixed solve (FP64 CG on
the matrix-free K, FP32 V-cycle with the matrix-free finest smoother)
reaches 1e-8 on the assembled FP64 residual, agrees with the all-FP64 solve
to 1e-6 and is bitwise reproducible."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    import torch
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


@pytest.fixture(scope="module")
def ibeam():
    from wnfea_amd.mesh.rcm import apply_rcm_to_model
    from wnfea_amd.mesh.solid_generator import generate_c3d10_ibeam
    from wnfea_amd.solver.solid_solver import MatrixFreeOperator, assemble_c3d10_bsr, build_amg
    model, meta = generate_c3d10_ibeam(n_layers=300, length=3.0)
    assert meta["n_dofs"] == 1054755
    model, _, _ = apply_rcm_to_model(model)
    system = assemble_c3d10_bsr(model, want_fp64_equilibrated=True)
    amg = build_amg(system, storage=("fp64", "fp32"), theta=0.1, theta_decay=0.4)
    op = MatrixFreeOperator(system)
    op.attach_smoother(amg)
    yield system, amg, op
    op.close()
    amg.close()


def test_matrix_free_operator_is_the_assembled_operator(ibeam):
    import torch
    system, _, op = ibeam
    assert op.affine
    g = torch.Generator(device="cuda").manual_seed(5)
    u = torch.randn(system.n_dofs, dtype=torch.float64, device="cuda", generator=g)
    v = torch.randn(system.n_dofs, dtype=torch.float64, device="cuda", generator=g)
    Ku, Kv = op.apply(u), op.apply(v)
    assert _rel(Ku, system.K.spmv(u)) <= 1e-13
    assert abs(float(v @ Ku) - float(u @ Kv)) <= 1e-12 * float(torch.linalg.norm(u) * torch.linalg.norm(Kv))


def test_mixed_solve_meets_tolerance_and_matches_fp64(ibeam):
    from wnfea_amd.solver.solid_solver import solve_mixed, true_residual
    system, amg, op = ibeam
    u, it, runs, rel, _ = solve_mixed(system, amg, tol=1e-8, amg_precision="fp32", operator=op)
    assert true_residual(system, u) <= 1e-8 and it <= 60
    u64, it64, _, _, _ = solve_mixed(system, amg, tol=1e-8, amg_precision="fp64", operator=op)
    assert true_residual(system, u64) <= 1e-8
    assert _rel(u, u64) <= 1e-6
    u2, it2, _, _, _ = solve_mixed(system, amg, tol=1e-8, amg_precision="fp32", operator=op)
    assert it2 == it and bool((u2 == u).all())
