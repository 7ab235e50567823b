"""Active-DOF bookkeeping, rigid couplings and the host PCG driver of the
nonlinear path, on the CPU (no device call)."""
import numpy as np
import pytest


def _mixed_model():
    """A C3D10 box clamped at x = 0 whose x = 0.4 face is driven through a
    rigid coupling by the first node of a 4-beam cantilever along +x (a
    member direction where the reference's linear beam frame is proper)."""
    from wnfea_amd.boundary.conditions import LoadDef, create_fixed_support
    from wnfea_amd.boundary.coupling import RigidCoupling
    from wnfea_amd.mesh.solid_generator import generate_c3d10_structured_block
    from wnfea_amd.model import PropertyAssignment
    from wnfea_amd.properties.materials import MaterialDef
    from wnfea_amd.properties.sections import create_solid_circle
    box, _ = generate_c3d10_structured_block(0.4, 0.1, 0.1, 4, 2, 2, tip_load_total=0.0, tets_per_hex=6,
                                             load_direction="z")
    nodes = np.asarray(box.mesh_nodes)
    nb = len(nodes)
    beam_nodes = np.array([[0.4 + 0.2 * k, 0.05, 0.05] for k in range(1, 5)])
    # the beam's first node is the master, placed at the centre of the x = 0.4 face
    model = box
    model.mesh_nodes = np.vstack([nodes, [[0.4, 0.05, 0.05]], beam_nodes])
    m = nb
    model.mesh_elements = np.array([[m + k, m + k + 1] for k in range(4)])
    model.materials["S"] = MaterialDef(name="S", youngs_modulus=2.1e11, poissons_ratio=0.3, yield_strength=2.5e8)
    sec = create_solid_circle("rod", 0.01)
    model.sections[sec.name] = sec
    model.element_properties = dict.fromkeys(range(4), PropertyAssignment("S", sec.name))
    face = np.flatnonzero(np.abs(nodes[:, 0] - 0.4) < 1e-12)
    far = np.flatnonzero(np.abs(nodes[:, 0]) < 1e-12)
    model.couplings = [RigidCoupling(master_node_id=m, slave_node_ids=face.tolist())]
    model.supports = [create_fixed_support(int(i), is_geometry_node=False) for i in far]
    model.loads = [LoadDef(node_id=m + 4, fy=-50.0, fz=20.0, mx=3.0, is_geometry_node=False)]
    return model, m, face


def test_coupling_matrices_give_rigid_motion():
    from wnfea_amd.boundary.coupling import RigidCoupling
    rng = np.random.default_rng(0)
    xm = rng.standard_normal(3)
    xs = rng.standard_normal((5, 3))
    T = RigidCoupling(0, list(range(5))).compute_coupling_matrices(xm, xs)
    t, th = rng.standard_normal(3), rng.standard_normal(3)
    for Ti, x in zip(T, xs):
        np.testing.assert_allclose(Ti @ np.concatenate([t, th]), t + np.cross(th, x - xm), rtol=1e-14, atol=1e-14)


def test_dof_manager_pure_beam_is_identity():
    from wnfea_amd.mesh.lattice import generate_beam_lattice
    from wnfea_amd.solver.dof_manager import DOFManager
    model, _ = generate_beam_lattice(2)
    d = DOFManager(model)
    n = len(model.mesh_nodes)
    assert d.total_active_dofs == 6 * n and not d.slave_nodes
    u = np.random.default_rng(1).standard_normal(6 * n)
    np.testing.assert_array_equal(d.expand_displacements(u), u)
    np.testing.assert_array_equal(d.condense_forces(u), u)
    assert d.node_dof_to_global[(3, 4)] == 22


def test_dof_manager_mixed_model():
    from wnfea_amd.solver.dof_manager import DOFManager
    model, m, face = _mixed_model()
    d = DOFManager(model)
    n = len(model.mesh_nodes)
    nb = m
    # beam nodes 6, pure solid nodes 3, coupled face nodes none
    assert d.total_active_dofs == 6 * 5 + 3 * (nb - len(face))
    assert all(len(d.active_dofs_per_node[i]) == 0 for i in face)
    assert len(d.active_dofs_per_node[m]) == 6 and len(d.active_dofs_per_node[0 if 0 not in face else nb - 1]) == 3
    rng = np.random.default_rng(2)
    ua = rng.standard_normal(d.total_active_dofs)
    f = rng.standard_normal(6 * n)
    uf = d.expand_displacements(ua)
    # the two maps are transposes of each other (virtual work) and of the projection
    assert abs(d.condense_forces(f) @ ua - f @ uf) <= 1e-12 * abs(f @ uf)
    np.testing.assert_allclose(d.projection() @ ua, uf, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(d.build_projection_matrix() @ ua, uf, rtol=1e-14, atol=1e-14)
    # slaves follow the master rigidly
    U = uf.reshape(-1, 6)
    X = np.asarray(model.mesh_nodes)
    np.testing.assert_allclose(U[face, :3], U[m, :3] + np.cross(U[m, 3:], X[face] - X[m]), rtol=1e-13, atol=1e-13)
    assert np.all(U[face, 3:] == 0.0)


def test_loads_and_constraints_in_active_numbering():
    from wnfea_amd.solver.dof_manager import DOFManager
    from wnfea_amd.solver.residual import build_external_force_vector, get_boundary_constraints
    model, m, _ = _mixed_model()
    d = DOFManager(model)
    F = build_external_force_vector(model)
    assert F.shape == (6 * len(model.mesh_nodes),)
    np.testing.assert_array_equal(F[6 * (m + 4):6 * (m + 5)], [0.0, -50.0, 20.0, 3.0, 0.0, 0.0])
    Fa = build_external_force_vector(model, d)
    assert Fa.shape == (d.total_active_dofs,) and abs(Fa.sum() - F.sum()) < 1e-12
    c, v = get_boundary_constraints(model, d)
    assert len(c) == 3 * len(model.supports) and np.all(v == 0.0)  # solid nodes: translations only
    c6, _ = get_boundary_constraints(model)
    assert len(c6) == 6 * len(model.supports)


def test_pcg_driver_on_a_small_spd_system():
    import scipy.sparse as sp
    from scipy.sparse.linalg import aslinearoperator, LinearOperator
    from wnfea_amd.solver.nonlinear_solver import pcg_solve
    n = 300
    A = sp.diags([-np.ones(n - 1), 2.5 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsr()
    b = np.random.default_rng(3).standard_normal(n)
    dinv = 1.0 / A.diagonal()
    M = LinearOperator((n, n), matvec=lambda r: dinv * r)
    stats = {}
    x, it, rn = pcg_solve(aslinearoperator(A), b, M_prec=M, tol=1e-10, max_iter=500, stats=stats)
    assert np.linalg.norm(b - A @ x) <= 1e-10 * np.linalg.norm(b) * 1.01 and rn == pytest.approx(
        np.linalg.norm(b - A @ x), rel=1e-6, abs=1e-12)
    assert stats["fp32_iters"] == it and stats["fp16_iters"] == 0  # one before the loop, none after the last
    x32, _, _ = pcg_solve(aslinearoperator(A), b, M_prec=M, tol=1e-5, working_dtype=np.float32)
    assert x32.dtype == np.float64 and np.linalg.norm(x32 - x) <= 1e-4 * np.linalg.norm(x)
