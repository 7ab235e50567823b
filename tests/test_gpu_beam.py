"""Co-rotational beam kernels and the device Newton solver against the numpy
oracle of the reference's beam path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def lattice():
    from wnfea_amd.mesh.lattice import beam_props, generate_beam_lattice
    from wnfea_amd.mesh.rcm import apply_rcm_to_model
    model, _ = generate_beam_lattice(5, lateral_sigma=500.0, compression_mean=2.0e3, compression_sigma=500.0)
    model, _, _ = apply_rcm_to_model(model)
    return model, beam_props(model)


def test_beam_forces_match_oracle(lattice):
    import torch
    from oracle import beam_oracle as B
    from wnfea_amd.solver.beam_solver import BeamSystem
    model, props = lattice
    bs = BeamSystem(model)
    rng = np.random.default_rng(0)
    for amp in (1e-6, 1e-3, 5e-2):
        U = amp * rng.standard_normal(6 * len(model.mesh_nodes))
        F = bs.internal_forces(torch.as_tensor(U, device="cuda")).cpu().numpy()
        Fo = B.corot_forces(model.mesh_nodes, model.mesh_elements, props, U)
        assert _rel(F, Fo) <= 1e-12, amp


def test_tangent_and_linear_stiffness_match_oracle(lattice):
    import torch
    from oracle import beam_oracle as B
    from wnfea_amd.solver.beam_solver import BeamSystem
    model, props = lattice
    bs = BeamSystem(model)
    U = 1e-3 * np.random.default_rng(1).standard_normal(6 * len(model.mesh_nodes))
    Kt = bs.element_tangents(torch.as_tensor(U, device="cuda")).cpu().numpy()
    Kto = B.fd_tangent(model.mesh_nodes, model.mesh_elements, props, U)
    assert np.abs(Kt - Kto).max() <= 1e-6 * np.abs(Kto).max()
    # assembled symmetric tangent with the Dirichlet rule
    bs.assemble_tangent(torch.as_tensor(U, device="cuda"))
    A = bs.system.K.to_scipy_bsr().tocsr()
    Ao = B.assemble(len(model.mesh_nodes), model.mesh_elements, Kto, model.fixed_dof_mask(6))
    assert abs(A - Ao).max() <= 1e-6 * abs(Ao).max()
    # the reference's linear stiffness, frame rules included
    bs.assemble_linear()
    Al = bs.system.K.to_scipy_bsr().tocsr()
    Alo = B.assemble(len(model.mesh_nodes), model.mesh_elements,
                     B.ke_linear(model.mesh_nodes, model.mesh_elements, props), model.fixed_dof_mask(6), sym=False)
    assert abs(Al - Alo).max() <= 1e-12 * abs(Alo).max()


def test_newton_matches_oracle(lattice):
    from oracle import beam_oracle as B
    from wnfea_amd.solver.beam_solver import BeamSystem
    model, props = lattice
    fixed = model.fixed_dof_mask(6)
    F = model.load_vector(6)
    F[fixed] = 0.0
    Uo, _ = B.newton(model.mesh_nodes, model.mesh_elements, props, F, fixed, n_steps=4, tol_rel=1e-10)
    bs = BeamSystem(model)
    try:
        U, st = bs.solve_newton(n_load_steps=4, tol_rel=1e-10, tol_abs=1e-9)
    finally:
        bs.close()
    assert st["final_residual"] <= 1e-6 * np.linalg.norm(F)
    assert _rel(U.cpu().numpy(), Uo) <= 1e-6
    assert np.linalg.norm(Uo) > 0
