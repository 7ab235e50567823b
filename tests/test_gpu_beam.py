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


def test_consistent_tangent_is_the_derivative_of_the_forces(lattice):
    """The default element tangent (h_rel = 0) is the exact derivative of the
    co-rotational forces: at a deformed state it equals the oracle's central
    differences extrapolated to h -> 0 (Richardson on h = 1e-4 and 5e-5) to
    1e-8, and at U = 0 (no axial force, so no geometric stiffness) the
    reference's linear stiffness of every member whose frame both codes build
    alike (+x and +z chords, right-handed) to round-off."""
    import torch
    from oracle import beam_oracle as B
    from wnfea_amd.solver.beam_solver import BeamSystem
    model, props = lattice
    bs = BeamSystem(model)
    assert bs.h_rel == 0.0
    U = 1e-3 * np.random.default_rng(2).standard_normal(6 * len(model.mesh_nodes))
    Kt = bs.element_tangents(torch.as_tensor(U, device="cuda")).cpu().numpy()
    K1 = B.fd_tangent(model.mesh_nodes, model.mesh_elements, props, U, h_rel=1e-4)
    K2 = B.fd_tangent(model.mesh_nodes, model.mesh_elements, props, U, h_rel=5e-5)
    R = (4.0 * K2 - K1) / 3.0
    assert np.abs(Kt - R).max() <= 1e-8 * np.abs(R).max()
    K0 = bs.element_tangents(torch.zeros(6 * len(model.mesh_nodes), dtype=torch.float64, device="cuda")).cpu().numpy()
    conn = np.asarray(model.mesh_elements)
    d = model.mesh_nodes[conn[:, 1]] - model.mesh_nodes[conn[:, 0]]
    sel = ((d[:, 0] > 0) & (np.abs(d[:, 1]) < 1e-12) & (np.abs(d[:, 2]) < 1e-12)) | \
          ((d[:, 2] > 0) & (np.abs(d[:, 0]) < 1e-12) & (np.abs(d[:, 1]) < 1e-12))
    assert sel.sum() > len(conn) // 3
    Kl = B.ke_linear(model.mesh_nodes, conn[sel], props[sel])
    assert np.abs(K0[sel] - Kl).max() <= 1e-12 * np.abs(Kl).max()
    # central differences stay available (h_rel > 0) and agree to O(h^2)
    bs.h_rel = 1e-6
    Kfd = bs.element_tangents(torch.as_tensor(U, device="cuda")).cpu().numpy()
    assert np.abs(Kfd - Kt).max() <= 1e-6 * np.abs(Kt).max()


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


def test_element_api_matches_oracle():
    """The single-element functions of the reference's beam API."""
    from oracle import beam_oracle as B
    from wnfea_amd.solver.corotational_beam import (build_corotational_projector, compute_corotational_element_forces,
                                                    compute_element_triad)
    rng = np.random.default_rng(5)
    E, G, A, Iy, Iz, J = 7e10, 2.6e10, 1.2e-3, 1.1e-7, 1.3e-7, 2.4e-7
    for X2 in (np.array([0.6, 0.2, 0.9]), np.array([0.0, 1.0, 0.0]), np.array([0.0, 0.0, -0.7])):
        X1 = np.zeros(3)
        u = 1e-2 * rng.standard_normal(12)
        f = compute_corotational_element_forces(X1, X2, u, E, G, A, Iy, Iz, J)
        fo = B.corot_element(X1, X2, u, E, G, A, Iy, Iz, J)
        assert _rel(f, fo) <= 1e-10
        r1, r2, r3, L, L0 = compute_element_triad(X1, X2, u)
        Rm = np.array([r1, r2, r3])
        assert np.abs(Rm @ Rm.T - np.eye(3)).max() <= 1e-14
        d = X2 + u[6:9] - u[:3]
        assert abs(L - np.linalg.norm(d)) <= 1e-14 and np.allclose(r1, d / L)
        Bm = build_corotational_projector(r1, r2, r3, L)
        t = rng.standard_normal(3)
        assert np.abs(Bm @ np.concatenate([t, [0, 0, 0], t, [0, 0, 0]])).max() <= 1e-14  # rigid translation


def test_residual_and_jfnk_operator_match_oracle(lattice):
    """Equilibrium residual (device forces, host bookkeeping) and the
    Jacobian-free tangent action against the oracle's forces and tangent."""
    from oracle import beam_oracle as B
    from wnfea_amd.solver.jfnk_operator import MatrixFreeJFNKOperator
    from wnfea_amd.solver.residual import build_external_force_vector, compute_equilibrium_residual
    model, props = lattice
    n6 = 6 * len(model.mesh_nodes)
    fixed = model.fixed_dof_mask(6)
    rng = np.random.default_rng(6)
    U = 1e-3 * rng.standard_normal(n6)
    U[fixed] = 0.0
    F = build_external_force_vector(model)
    R = compute_equilibrium_residual(model, U, F, 0.5)
    Ro = B.corot_forces(model.mesh_nodes, model.mesh_elements, props, U) - 0.5 * F
    Ro[fixed] = U[fixed]
    assert _rel(R, Ro) <= 1e-12
    Jop = MatrixFreeJFNKOperator(model, U, F, load_factor=0.5)
    Kt = B.assemble(len(model.mesh_nodes), model.mesh_elements,
                    B.fd_tangent(model.mesh_nodes, model.mesh_elements, props, U), fixed, sym=False)
    for _ in range(2):
        v = rng.standard_normal(n6)
        assert _rel(Jop.matvec(v), Kt @ v) <= 1e-5


def test_host_pcg_with_jfnk_and_reference_amg(lattice):
    import scipy.sparse.linalg as sla
    from oracle import beam_oracle as B
    from wnfea_amd.solver.amg_preconditioner import BlockBeamAMGPreconditioner
    from wnfea_amd.solver.jfnk_operator import MatrixFreeJFNKOperator
    from wnfea_amd.solver.nonlinear_solver import pcg_solve
    from wnfea_amd.solver.residual import build_external_force_vector
    model, props = lattice
    F = build_external_force_vector(model)
    fixed = model.fixed_dof_mask(6)
    F[fixed] = 0.0
    U = np.zeros_like(F)
    Jop = MatrixFreeJFNKOperator(model, U, F, load_factor=1.0)
    amg = BlockBeamAMGPreconditioner(model, working_dtype=np.float32)
    amg.enable_tri_precision(1e-2)
    stats = {}
    x, it, rn = pcg_solve(Jop, -Jop.R_current, M_prec=amg, tol=1e-6, max_iter=200, stats=stats)
    assert rn <= 1e-6 * np.linalg.norm(Jop.R_current) and it < 200
    assert stats["fp16_iters"] >= 1 and stats["fp32_iters"] >= 1
    # the co-rotational tangent at U = 0 (the reference's linear beam Ke uses
    # another frame for +y members, so it is not this operator)
    Kt = B.assemble(len(model.mesh_nodes), model.mesh_elements,
                    B.fd_tangent(model.mesh_nodes, model.mesh_elements, props, U), fixed, sym=False)
    x_ref = sla.spsolve(Kt.tocsc(), -Jop.R_current)
    assert _rel(x, x_ref) <= 1e-4


def test_mixed_beam_solid_model_is_consistent():
    """Beam + C3D10 box joined by a rigid coupling.  The device residual of
    the active-DOF system equals the oracle's (co-rotational beams + C3D10
    internal forces, condensed through the DOF manager), and its central
    difference at U = 0 is the linear operator T^T K T of the 6-DOF assembly.
    The linear solution itself is not a nonlinear equilibrium: the rod's
    axial stiffness turns its O(v^2/L) shortening under the lateral load into
    axial forces of the size of the load on the beam nodes' ux (the oracle
    gives 0.40 there against |F| = 0.54, and 1e-3 on the other DOFs)."""
    import scipy.sparse.linalg as sla
    from oracle import beam_oracle as B
    from oracle import oracle as O
    from test_cpu_dof import _mixed_model
    from wnfea_amd.mesh.lattice import beam_props
    from wnfea_amd.solver.beam_solver import assemble_global_system_sparse
    from wnfea_amd.solver.dof_manager import DOFManager
    from wnfea_amd.solver.residual import (build_external_force_vector, compute_equilibrium_residual,
                                           get_boundary_constraints)
    model, m, face = _mixed_model()
    for ld in model.loads:
        ld.fy, ld.fz, ld.mx = 1e-2 * ld.fy, 1e-2 * ld.fz, 1e-2 * ld.mx
    d = DOFManager(model)
    K, F = assemble_global_system_sparse(model)
    T = d.projection()
    Ka = (T.T @ K @ T).tocsc()
    Fa = build_external_force_vector(model, d)
    ua = sla.spsolve(Ka, Fa)
    assert np.linalg.norm(Ka @ ua - Fa) <= 1e-10 * np.linalg.norm(Fa)
    U = d.expand_displacements(ua).reshape(-1, 6)
    assert np.abs(U[m + 4, :3]).max() > 0 and np.abs(U[face, :3]).max() > 0
    # device residual == oracle residual at the linear solution
    X = np.asarray(model.mesh_nodes, dtype=np.float64)
    E, nu = model.element_moduli()
    f_or = (B.corot_forces(X, model.mesh_elements, beam_props(model), U.ravel())
            + O.fint_global(X, np.asarray(model.solid_elements), np.column_stack([E, nu]), U.ravel(), stride=6))
    c, _ = get_boundary_constraints(model, d)
    R_or = d.condense_forces(f_or) - Fa
    R_or[c] = ua[c]
    R = compute_equilibrium_residual(model, ua, Fa, 1.0, dof_mgr=d)
    assert _rel(R, R_or) <= 1e-10
    beam_ux = np.array([d.node_dof_to_global[(m + k, 0)] for k in range(5)])
    rest = np.setdiff1d(np.arange(len(R)), beam_ux)
    assert np.linalg.norm(R[rest]) <= 1e-2 * np.linalg.norm(Fa) < np.linalg.norm(R[beam_ux])
    # the residual's central difference at 0 is the assembled operator
    v = np.random.default_rng(5).standard_normal(len(ua))
    v[c] = 0.0
    h = 1e-7 / np.abs(v).max()
    dR = (compute_equilibrium_residual(model, h * v, 0 * Fa, 1.0, dof_mgr=d)
          - compute_equilibrium_residual(model, -h * v, 0 * Fa, 1.0, dof_mgr=d)) / (2 * h)
    Kv = Ka @ v
    keep = np.setdiff1d(np.arange(len(v)), c)
    assert _rel(dR[keep], Kv[keep]) <= 1e-6
