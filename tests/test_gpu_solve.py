"""Device solvers against the CPU FP64 oracle on C0 (19 683 DOF)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c0_system(c0_model):
    from wnfea_amd.solver.solid_solver import assemble_c3d10_bsr
    _, mr, _ = c0_model
    return assemble_c3d10_bsr(mr, want_fp64_equilibrated=True)


@pytest.fixture(scope="module")
def c0_amg(c0_system):
    from wnfea_amd.solver.solid_solver import build_amg
    return build_amg(c0_system)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_jacobi_pcg_fp64(c0_system, c0_oracle):
    from wnfea_amd.solver.solid_solver import pcg_device, true_residual
    _, _, u_ref = c0_oracle
    x, it, rel, _ = pcg_device(c0_system, rtol=1e-10, max_iter=20000, preconditioner="jacobi", precision="fp64")
    assert rel <= 1e-10  # recursive residual; the true one drifts above it as in any CG
    assert true_residual(c0_system, x) <= 1e-8
    assert _rel(x.cpu().numpy(), u_ref) <= 1e-7
    assert 100 < it < 20000


@pytest.mark.parametrize("precision", ["fp64", "fp64_equilibrated"])
def test_amg_pcg_fp64(c0_system, c0_amg, c0_oracle, precision):
    from wnfea_amd.solver.solid_solver import pcg_device, true_residual
    _, _, u_ref = c0_oracle
    x, it, rel, _ = pcg_device(c0_system, rtol=1e-10, max_iter=500, preconditioner="amg", precision=precision,
                               amg=c0_amg, amg_precision="fp64")
    assert it < 100
    assert true_residual(c0_system, x) <= 1e-9
    assert _rel(x.cpu().numpy(), u_ref) <= 1e-8


@pytest.mark.parametrize("amg_precision", ["fp32", "fp16", "fp16_fine", "fp16_coarse", "fp16_smooth"])
def test_mixed_precision_preconditioner(c0_system, c0_amg, c0_oracle, amg_precision):
    """FP64 CG with an FP32/FP16 V-cycle: FP64-verified 1e-8 and the 1e-6 mixed-mode displacement bar
    (fp16_smooth: FP16 smoothing steps, FP32 defects, the flexible beta)."""
    from wnfea_amd.solver.solid_solver import pcg_device, solve_mixed, true_residual
    _, _, u_ref = c0_oracle
    u, it, steps, rel, _ = solve_mixed(c0_system, c0_amg, tol=1e-8, amg_precision=amg_precision)
    assert rel <= 1e-8 and true_residual(c0_system, u) <= 1e-8
    assert _rel(u.cpu().numpy(), u_ref) <= 1e-6
    _, it64, _, _ = pcg_device(c0_system, rtol=0.5e-8, preconditioner="amg", amg=c0_amg, amg_precision="fp64")
    # an FP32 cycle is the FP64 one to ~1e-6; FP16 storage weakens the preconditioner
    assert it <= (it64 + 5 if amg_precision == "fp32" else 2 * it64 + 5)


@pytest.mark.parametrize("amg_precision", ["fp64", "fp32"])
def test_graph_replay_matches_host_launches(c0_system, c0_amg, amg_precision):
    """Replayed iteration graphs run the same kernels in the same order as host
    launches, so the iterates are bitwise identical; a cached graph is reused
    and a different chunk size captures a new one."""
    import torch
    from wnfea_amd.solver.solid_solver import pcg_device, set_graphs
    kw = dict(rtol=1e-9, max_iter=400, preconditioner="amg", precision="fp64", amg=c0_amg,
              amg_precision=amg_precision)
    try:
        set_graphs(False)
        x0, it0, rel0, _ = pcg_device(c0_system, check_every=8, **kw)
        set_graphs(True)
        x1, it1, rel1, _ = pcg_device(c0_system, check_every=8, **kw)
        x2, it2, rel2, _ = pcg_device(c0_system, check_every=8, **kw)  # replays the cached graph
        x3, it3, _, _ = pcg_device(c0_system, check_every=5, **kw)       # new capture, ragged last chunk
    finally:
        set_graphs(False)  # the default
    assert it0 == it1 == it2 == it3 and rel0 == rel1 == rel2
    assert torch.equal(x0, x1) and torch.equal(x1, x2) and torch.equal(x1, x3)
    # a capped run stops at exactly max_iter with and without graphs
    kw["rtol"] = 1e-30
    kw["max_iter"] = 13
    set_graphs(True)
    xa, ita, _, _ = pcg_device(c0_system, check_every=4, **kw)
    set_graphs(False)
    xb, itb, _, _ = pcg_device(c0_system, check_every=4, **kw)
    assert ita == itb == 13 and torch.equal(xa, xb)


def test_mixed_ir_amg(c0_system, c0_amg, c0_oracle):
    from wnfea_amd.solver.solid_solver import ir_device, true_residual
    _, _, u_ref = c0_oracle
    u, outer, inner, rel, _ = ir_device(c0_system, tol=1e-8, amg=c0_amg)
    assert rel <= 1e-8 and true_residual(c0_system, u) <= 1e-8
    assert _rel(u.cpu().numpy(), u_ref) <= 1e-6
    assert outer <= 12


def test_vcycle_fp32_vs_fp64(c0_system, c0_amg):
    import torch
    n = c0_system.n_dofs
    r = torch.as_tensor(np.random.default_rng(0).standard_normal(n), device=c0_system.K.device)
    z64 = c0_amg.vcycle(r)
    z32 = c0_amg.vcycle(r.float())
    z16 = c0_amg.vcycle(r.float(), storage="fp16")
    z16f = c0_amg.vcycle(r.float(), storage="fp16_fine")
    assert _rel(z32.double().cpu().numpy(), z64.cpu().numpy()) <= 1e-4
    assert _rel(z16.double().cpu().numpy(), z64.cpu().numpy()) <= 5e-2
    assert _rel(z16f.double().cpu().numpy(), z64.cpu().numpy()) <= 5e-2
    # symmetric positive preconditioner: r.Mr > 0 and x.My = y.Mx
    s = torch.as_tensor(np.random.default_rng(1).standard_normal(n), device=r.device)
    zs = c0_amg.vcycle(s)
    assert float(r @ z64) > 0
    assert abs(float(r @ zs) - float(s @ z64)) <= 1e-8 * abs(float(r @ z64))


def test_drop_in_solvers(c0_model, c0_oracle):
    from wnfea_amd.solver.solid_solver import solve_c3d10_pcg, solve_c3d10_system
    K, F, u_ref = c0_oracle
    u, it, rel, _ = solve_c3d10_pcg(K, F, rtol=1e-8, max_iter=20000, preconditioner="jacobi")
    assert rel < 1e-7 and _rel(u, u_ref) < 1e-6
    u2, it2, rel2, _ = solve_c3d10_pcg(K, F, rtol=1e-8, max_iter=20000, preconditioner="jacobi",
                                       precision="mixed_fp32")
    assert _rel(u2, u_ref) < 1e-5
    u3, st = solve_c3d10_system(K, F)  # auto -> direct-tolerance path
    # the direct solve's own FP64 residual is the bar for "direct tolerance"
    r_direct = np.linalg.norm(F - K @ u_ref) / np.linalg.norm(F)
    assert st["rel_res"] <= max(10 * r_direct, 1e-12)
    assert _rel(u3, u_ref) < 1e-9


def test_vcycle_matches_cpu_mirror(c0_amg):
    """FP64 device V-cycle against the numpy restatement of the same cycle on
    the same host hierarchy."""
    import torch
    from vcycle_mirror import vcycle
    n = c0_amg.levels[0].A.shape[0]
    r = np.random.default_rng(3).standard_normal(n)
    z_ref = vcycle(c0_amg.levels, c0_amg.coarse_inv, r, deg=c0_amg.level_degrees, ratio=c0_amg.ratio)
    z = c0_amg.vcycle(torch.as_tensor(r, device="cuda")).cpu().numpy()
    assert _rel(z, z_ref) <= 1e-12


def test_stress_field_vs_oracle(c0_model, c0_oracle):
    from oracle import oracle as O
    from wnfea_amd.solver.solid_solver import compute_c3d10_stress_field
    _, mr, _ = c0_model
    _, _, u_ref = c0_oracle
    cs, cv, nv = compute_c3d10_stress_field(mr, u_ref)
    E, nu = mr.element_moduli()
    cso, cvo, nvo = O.stress_field(mr.mesh_nodes, mr.solid_elements, E, nu, u_ref)
    assert np.abs(cs - cso).max() <= 1e-9 * np.abs(cso).max()
    assert np.abs(cv - cvo).max() <= 1e-9 * cvo.max()
    assert np.abs(nv - nvo).max() <= 1e-9 * nvo.max()


def test_vtu_export_stresses_match_oracle(c0_model, c0_oracle, tmp_path):
    """The .vtu written for the solved C0 box carries the oracle's element and
    nodal von Mises and element stress tensors (to the file's 9 digits)."""
    import copy
    import xml.etree.ElementTree as ET
    from oracle import oracle as O
    _, mr, _ = c0_model
    _, _, u_ref = c0_oracle
    m = copy.copy(mr)
    m.displacements = u_ref
    path = m.export_vtu(tmp_path / "c0.vtu")
    arrays = {da.get("Name"): np.array(da.text.split(), dtype=np.float64)
              for da in ET.parse(path).getroot().iter("DataArray")}
    E, nu = mr.element_moduli()
    cso, cvo, nvo = O.stress_field(mr.mesh_nodes, mr.solid_elements, E, nu, u_ref)
    np.testing.assert_allclose(arrays["VonMises"], cvo, rtol=1e-8, atol=1e-9 * cvo.max())
    np.testing.assert_allclose(arrays["VonMises_Nodal"], nvo, rtol=1e-8, atol=1e-9 * nvo.max())
    np.testing.assert_allclose(arrays["StressTensor"].reshape(-1, 6), cso, rtol=1e-8, atol=1e-9 * np.abs(cso).max())
    np.testing.assert_allclose(arrays["Displacement"], u_ref, rtol=1e-8, atol=1e-9 * np.abs(u_ref).max())


def test_gmsh_box_solves_like_generated_box(tmp_path):
    """A C3D10 box written as a gmsh file (type-11 node order) and read back
    solves to the same displacements as the generated box."""
    from test_cpu_io import _to_gmsh, _write
    from wnfea_amd.mesh.gmsh_reader import get_nodes_on_plane, load_msh
    from wnfea_amd.mesh.solid_generator import generate_c3d10_structured_block
    from wnfea_amd.solver.solid_solver import assemble_c3d10_sparse_3dof, solve_c3d10_system
    from wnfea_amd.boundary.conditions import LoadDef, create_fixed_support
    gen, _ = generate_c3d10_structured_block(0.6, 0.1, 0.1, 12, 2, 2, tip_load_total=-1000.0, tets_per_hex=6,
                                             load_direction="z")
    nodes, elems = np.asarray(gen.mesh_nodes), np.asarray(gen.solid_elements, dtype=np.int64)
    path = tmp_path / "box.msh"
    _, order = _write(path, nodes, _to_gmsh(nodes, elems), 4, False)
    n, e, m = load_msh(path, material_name="Steel")
    root = get_nodes_on_plane(n, "x", 0.0)
    tip = get_nodes_on_plane(n, "x", 0.6)
    m.supports = [create_fixed_support(int(i), is_geometry_node=False) for i in root]
    m.loads = [LoadDef(node_id=int(i), fz=-1000.0 / len(tip), is_geometry_node=False) for i in tip]
    K, F, _ = assemble_c3d10_sparse_3dof(m)
    u, _ = solve_c3d10_system(K, F, rtol=1e-10)
    Kg, Fg, _ = assemble_c3d10_sparse_3dof(gen)
    ug, _ = solve_c3d10_system(Kg, Fg, rtol=1e-10)
    u_back = u.reshape(-1, 3)
    np.testing.assert_allclose(u_back, ug.reshape(-1, 3)[order], rtol=0, atol=1e-8 * np.abs(ug).max())


@pytest.mark.parametrize("operator", ["matrix_free", "assembled"])
def test_many_mesh_batch_matches_oracle(operator):
    """Six uneven boxes solved as one block-diagonal system: every mesh meets
    1e-8 on its own FP64 residual and matches its own exact solution."""
    from scipy.sparse.linalg import spsolve
    from oracle import oracle as O
    from wnfea_amd.solver.batch import BatchedSolids, generate_box_batch
    batch = generate_box_batch(6, seed=1, dof_range=(2000, 12000))
    models = [b[0] for b in batch]
    bs = BatchedSolids(models, operator=operator)
    try:
        u, iters, per, _ = bs.solve(tol=1e-8)
        assert np.all(per <= 1e-8)
        for m, uk in zip(models, bs.split(u)):
            E, nu = m.element_moduli()
            K, F = O.assemble_csr(m.mesh_nodes, m.solid_elements, E, nu, m.fixed_dof_mask(3), m.load_vector(3))
            ref = spsolve(K.tocsc(), F)
            assert _rel(uk, ref) <= 1e-6
    finally:
        bs.close()


@pytest.mark.parametrize("group", [4, 8])
def test_multi_rhs_matches_oracle(c0_system, c0_amg, c0_oracle, group):
    """Ten load cases through the multi-column solver (one SpMM per iteration
    for a whole group): each column meets 1e-8 on its FP64 residual and
    matches the exact solution of its own load case."""
    import torch
    from scipy.sparse.linalg import splu
    from wnfea_amd.solver.solid_solver import solve_multi
    K, F, _ = c0_oracle
    rng = np.random.default_rng(5)
    B = rng.standard_normal((K.shape[0], 10)) * np.abs(F).max()
    B[c0_system.fixed_mask] = 0.0
    B[:, 3] = 0.0  # an empty load case converges at once
    U, iters, rels, _ = solve_multi(c0_system, c0_amg, torch.as_tensor(B, device="cuda"), tol=1e-8, group=group)
    lu = splu(K.tocsc())
    Uh = U.cpu().numpy()
    for k in range(10):
        if k == 3:
            assert np.all(Uh[:, k] == 0.0) and iters[k] == 0
            continue
        ref = lu.solve(B[:, k])
        assert rels[k] <= 1e-8 and _rel(Uh[:, k], ref) <= 1e-6


def test_spmm_matches_spmv(c0_system):
    import torch
    from wnfea_amd import _lib
    L = _lib.load()
    n = c0_system.n_dofs
    X = torch.randn((n, 8), dtype=torch.float64, device="cuda")
    Y = torch.empty_like(X)
    _lib.check(L.wnfea_bsr_spmm(c0_system.K.handle, 0, 8, _lib.dptr(X), _lib.dptr(Y), _lib.stream_of()), "spmm")
    for k in range(8):
        y = c0_system.K.spmv(X[:, k].contiguous())
        assert float(torch.linalg.norm(Y[:, k] - y) / torch.linalg.norm(y)) <= 1e-14


@pytest.mark.parametrize("policy", ["MIXED_FP32", "TRI_PRECISION", "FULL_FP64"])
def test_precision_policies(c0_system, c0_amg, c0_oracle, policy):
    """The reference's presets (mixed_precision.py:96-122) on the device
    solver: FP64-verified 1e-8 and the displacement bars."""
    from wnfea_amd.solver import mixed_precision as mp
    _, _, u_ref = c0_oracle
    u, it, rel, _ = mp.solve_with_policy(c0_system, c0_amg, getattr(mp, policy), tol=1e-8)
    assert rel <= 1e-8
    assert _rel(u.cpu().numpy(), u_ref) <= 1e-6


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_vcycle_low_precision_matches_rounded_mirror(c0_amg, storage):
    """The FP32/FP16-stored device V-cycle equals the numpy cycle run on the
    same hierarchy with every operator (A, P, R on all levels) rounded to the
    storage type; the remaining difference is FP32 arithmetic."""
    import copy
    import scipy.sparse as sp
    import torch
    from vcycle_mirror import vcycle
    dt = np.float16 if storage == "fp16" else np.float32
    levels = copy.deepcopy(c0_amg.levels)

    def rnd(M):
        M = sp.csr_matrix(M).copy()
        M.data = M.data.astype(dt).astype(np.float64)
        return M

    for lv in levels[:-1]:
        lv.A = rnd(lv.A)
        lv.P, lv.R = rnd(lv.P), rnd(lv.R)
    n = levels[0].A.shape[0]
    r = np.random.default_rng(4).standard_normal(n)
    z_ref = vcycle(levels, c0_amg.coarse_inv, r, deg=c0_amg.level_degrees, ratio=c0_amg.ratio)
    z = c0_amg.vcycle(torch.as_tensor(r, device="cuda").float(), storage=storage).double().cpu().numpy()
    assert _rel(z, z_ref) <= 2e-5


@pytest.mark.parametrize("storage", ["fp64", "fp32"])
def test_coarse_only_cycle_is_inverse_product(storage):
    """A hierarchy of one level (1 575 DOF, below the coarse size): the cycle
    is the packed symmetric tile product of the exact inverse, so it must equal
    inv @ r (FP64: to round-off; FP32: against the inverse rounded to FP32,
    the result rounded once).  1 575 = 24 * 64 + 39 exercises a partial last
    tile and off-diagonal tiles on both sides."""
    import torch
    from wnfea_amd.mesh.rcm import apply_rcm_to_model
    from wnfea_amd.mesh.solid_generator import generate_c3d10_structured_block
    from wnfea_amd.solver.solid_solver import assemble_c3d10_bsr, build_amg
    m, _ = generate_c3d10_structured_block(1.0, 0.1, 0.1, 10, 2, 2, tip_load_total=-1000.0, tets_per_hex=6,
                                           load_direction="z")
    mr, _, _ = apply_rcm_to_model(m)
    system = assemble_c3d10_bsr(mr, want_fp64_equilibrated=True)
    amg = build_amg(system, storage=("fp64", "fp32"))
    assert len(amg.levels) == 1
    inv = amg.coarse_inv
    n = inv.shape[0]
    assert n == 1575 and n % 64 != 0
    r = np.random.default_rng(7).standard_normal(n)
    if storage == "fp64":
        z = amg.vcycle(torch.as_tensor(r, device="cuda")).cpu().numpy()
        assert _rel(z, inv @ r) <= 1e-13
    else:
        r32 = r.astype(np.float32)
        z_ref = inv.astype(np.float32).astype(np.float64) @ r32.astype(np.float64)
        z = amg.vcycle(torch.as_tensor(r32, device="cuda")).double().cpu().numpy()
        assert np.max(np.abs(z - z_ref) / np.maximum(np.abs(z_ref), 1e-30 + np.abs(z_ref).max() * 1e-6)) <= 2e-7
    amg.close()


def test_matrix_free_operator_matches_oracle_matrix(c0_system, c0_oracle):
    """The matrix-free K (quadrature data + on-the-fly B^T D B, Dirichlet rule
    of the assembled matrix) applies the oracle's assembled CSR K, and its
    residual equals the assembled one."""
    import torch
    from wnfea_amd.solver.solid_solver import MatrixFreeOperator, true_residual
    K, F, u_ref = c0_oracle
    op = MatrixFreeOperator(c0_system)
    rng = np.random.default_rng(11)
    for _ in range(2):
        u = rng.standard_normal(c0_system.n_dofs)
        y = op.apply(torch.as_tensor(u, device="cuda")).cpu().numpy()
        y_ref = K @ u
        assert _rel(y, y_ref) <= 1e-13
    ud = torch.as_tensor(u_ref + 1e-6 * rng.standard_normal(len(u_ref)) * np.abs(u_ref).max(), device="cuda")
    assert abs(op.residual(ud) - true_residual(c0_system, ud)) <= 1e-10 * true_residual(c0_system, ud)
    op.close()


@pytest.mark.parametrize("amg_precision", ["fp64", "fp32"])
def test_matrix_free_solve_matches_oracle(c0_system, c0_amg, c0_oracle, amg_precision):
    from wnfea_amd.solver.solid_solver import MatrixFreeOperator, solve_mixed, true_residual
    _, _, u_ref = c0_oracle
    op = MatrixFreeOperator(c0_system)
    u, it, steps, rel, _ = solve_mixed(c0_system, c0_amg, tol=1e-8, amg_precision=amg_precision, operator=op)
    assert rel <= 1e-8 and true_residual(c0_system, u) <= 1e-8
    assert _rel(u.cpu().numpy(), u_ref) <= (1e-8 if amg_precision == "fp64" else 1e-6)
    u2, it2, _, _, _ = solve_mixed(c0_system, c0_amg, tol=1e-8, amg_precision=amg_precision)
    assert abs(it - it2) <= 2
    op.close()


def test_fp16_smoothing_cycle_is_the_fp32_cycle_with_rounded_smoother(c0_amg):
    """The nonsymmetric cycle (FP16 smoothing steps, FP32 defects and
    transfers) equals the numpy cycle whose smoother operators are rounded to
    FP16 while defects use the FP32-rounded operators; it differs from the
    symmetric FP32 cycle by O(FP16 rounding)."""
    import copy
    import scipy.sparse as sp
    import torch
    from vcycle_mirror import vcycle
    levels = copy.deepcopy(c0_amg.levels)

    def rnd(M, dt):
        M = sp.csr_matrix(M).copy()
        M.data = M.data.astype(dt).astype(np.float64)
        return M

    for lv in levels[:-1]:
        lv.A_smooth = rnd(lv.A, np.float16)
        lv.A = rnd(lv.A, np.float32)
        lv.P, lv.R = rnd(lv.P, np.float32), rnd(lv.R, np.float32)
    n = levels[0].A.shape[0]
    r = np.random.default_rng(9).standard_normal(n)
    z_ref = vcycle(levels, c0_amg.coarse_inv, r, deg=c0_amg.level_degrees, ratio=c0_amg.ratio)
    rt = torch.as_tensor(r, device="cuda").float()
    z = c0_amg.vcycle(rt, storage="fp16_smooth").double().cpu().numpy()
    assert _rel(z, z_ref) <= 2e-5
    z32 = c0_amg.vcycle(rt, storage="fp32").double().cpu().numpy()
    assert 1e-6 <= _rel(z, z32) <= 2e-2


def test_matrix_free_smoothing_cycle_and_solve(c0_system, c0_oracle):
    """Finest-level smoothing steps and defect on the matrix-free operator
    (diag(s) K diag(s), the same exact operator as the assembled S K S): the
    FP64 cycle is still the numpy mirror's cycle, the FP32 one within FP32
    arithmetic of it, solves keep their accuracy and iteration counts, and
    detaching restores the assembled cycle bit for bit."""
    import torch
    from vcycle_mirror import vcycle
    from wnfea_amd.solver.solid_solver import MatrixFreeOperator, build_amg, solve_mixed, true_residual
    _, _, u_ref = c0_oracle
    amg = build_amg(c0_system, storage=("fp64", "fp32", "fp16"))
    n = amg.levels[0].A.shape[0]
    r = np.random.default_rng(12).standard_normal(n)
    z_ref = vcycle(amg.levels, amg.coarse_inv, r, deg=amg.level_degrees, ratio=amg.ratio)
    rd = torch.as_tensor(r, device="cuda")
    z64_asm = amg.vcycle(rd).cpu().numpy()
    z32_asm = amg.vcycle(rd.float()).double().cpu().numpy()
    op = MatrixFreeOperator(c0_system)
    op.attach_smoother(amg)
    assert amg.matrix_free_smoothing
    z64 = amg.vcycle(rd).cpu().numpy()
    assert _rel(z64, z_ref) <= 1e-12 and not np.array_equal(z64, z64_asm)
    z32 = amg.vcycle(rd.float()).double().cpu().numpy()
    assert _rel(z32, z_ref) <= 1e-4 and _rel(z32, z32_asm) <= 1e-4
    z16s = amg.vcycle(rd.float(), storage="fp16_smooth").double().cpu().numpy()
    assert _rel(z16s, z_ref) <= 2e-2
    its = {}
    for prec, tol_u in (("fp64", 1e-8), ("fp32", 1e-6), ("fp16_smooth", 1e-6)):
        u, it, _, rel, _ = solve_mixed(c0_system, amg, tol=1e-8, amg_precision=prec, operator=op)
        assert rel <= 1e-8 and true_residual(c0_system, u) <= 1e-8, prec
        assert _rel(u.cpu().numpy(), u_ref) <= tol_u, prec
        its[prec] = it
    op.close()  # detaches
    assert not amg.matrix_free_smoothing
    np.testing.assert_array_equal(amg.vcycle(rd).cpu().numpy(), z64_asm)
    for prec in ("fp64", "fp32"):
        _, it2, _, _, _ = solve_mixed(c0_system, amg, tol=1e-8, amg_precision=prec)
        assert abs(its[prec] - it2) <= 2, (prec, its[prec], it2)
    amg.close()


def test_matrix_free_operator_curved_and_affine_elements():
    """Generated meshes are affine (one inverse Jacobian per element); moving
    the midside nodes off the edge midpoints makes the elements curved (J
    varies inside them) and the operator keeps data per Gauss point.  Both
    apply the oracle's assembled K, and the matrix-free smoother on the curved
    mesh still gives the mirror's cycle."""
    import torch
    from oracle import oracle as O
    from vcycle_mirror import vcycle
    from wnfea_amd.mesh.rcm import apply_rcm_to_model
    from wnfea_amd.mesh.solid_generator import generate_c3d10_structured_block
    from wnfea_amd.solver.solid_solver import MatrixFreeOperator, assemble_c3d10_bsr, build_amg
    rng = np.random.default_rng(13)
    for curved in (False, True):
        m, _ = generate_c3d10_structured_block(0.4, 0.1, 0.1, 8, 2, 2, tip_load_total=-1000.0, tets_per_hex=6,
                                               load_direction="z")
        if curved:
            nodes = np.array(m.mesh_nodes, dtype=np.float64)
            mid = np.unique(np.asarray(m.solid_elements)[:, 4:].ravel())
            nodes[mid] += 0.001 * rng.standard_normal((len(mid), 3))  # 2 % of the 0.05 m element edge
            m.mesh_nodes = nodes
        mr, _, _ = apply_rcm_to_model(m)
        system = assemble_c3d10_bsr(mr, want_fp64_equilibrated=True)
        op = MatrixFreeOperator(system)
        assert op.affine == (not curved)
        E, nu = mr.element_moduli()
        K, _ = O.assemble_csr(mr.mesh_nodes, mr.solid_elements, E, nu, mr.fixed_dof_mask(3), mr.load_vector(3))
        for _ in range(2):
            u = rng.standard_normal(system.n_dofs)
            y = op.apply(torch.as_tensor(u, device="cuda")).cpu().numpy()
            assert _rel(y, K @ u) <= 1e-13, curved
        amg = build_amg(system, storage=("fp64", "fp32"))
        op.attach_smoother(amg)
        r = rng.standard_normal(system.n_dofs)
        z_ref = vcycle(amg.levels, amg.coarse_inv, r, deg=amg.level_degrees, ratio=amg.ratio)
        z = amg.vcycle(torch.as_tensor(r, device="cuda")).cpu().numpy()
        assert _rel(z, z_ref) <= 1e-12, curved
        z32 = amg.vcycle(torch.as_tensor(r, device="cuda").float()).double().cpu().numpy()
        assert _rel(z32, z_ref) <= 1e-4, curved
        op.close()
        amg.close()


def test_graph_replay_with_matrix_free_operators_after_reattach(c0_system):
    """Replayed iteration graphs that contain the matrix-free kernels are
    keyed on every buffer those kernels take: re-attaching the smoother (new
    scaling and quadrature copies) captures a new graph, and every result is
    bitwise the host-launch result."""
    import torch
    from wnfea_amd.solver.solid_solver import MatrixFreeOperator, build_amg, set_graphs, solve_mixed
    amg = build_amg(c0_system, storage=("fp64", "fp32"))
    op = MatrixFreeOperator(c0_system)
    try:
        op.attach_smoother(amg)
        set_graphs(False)
        x0, it0, _, _, _ = solve_mixed(c0_system, amg, tol=1e-8, operator=op, check_every=4)
        set_graphs(True)
        x1, it1, _, _, _ = solve_mixed(c0_system, amg, tol=1e-8, operator=op, check_every=4)
        op.attach_smoother(amg)  # fresh copies
        x2, it2, _, _, _ = solve_mixed(c0_system, amg, tol=1e-8, operator=op, check_every=4)
        assert it0 == it1 == it2
        assert torch.equal(x0, x1) and torch.equal(x0, x2)
    finally:
        set_graphs(False)
        op.close()
        amg.close()


def test_cycle_launch_shape_autotune(c0_system, tmp_path, monkeypatch):
    """The V-cycle's launch shapes are timed on whole cycles once the finest
    level is matrix-free: the pick is one of the candidates, it is cached, and
    every shape gives the same cycle up to FP32 rounding (the shapes only
    change the order of each row's sum)."""
    import torch
    from wnfea_amd.solver.autotune import autotune_cycle
    from wnfea_amd.solver.solid_solver import MatrixFreeOperator, build_amg
    monkeypatch.setenv("WNFEA_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    amg = build_amg(c0_system, storage=("fp64", "fp32"), autotune=False, waves=1, unroll=1, xcd_remap=True)
    op = MatrixFreeOperator(c0_system)
    try:
        op.attach_smoother(amg)  # built without autotuning: the shapes stay as given
        r = torch.as_tensor(np.random.default_rng(21).standard_normal(c0_system.n_dofs), device="cuda").float()
        z_ref = amg.vcycle(r).double().cpu().numpy()
        c = autotune_cycle(amg, "fp32", reps=3)
        assert not c["cached"] and (c["waves"], c["unroll"]) in {(w, u) for w in (1, 2, 4) for u in (1, 2)}
        assert len(c["table"]) == 6 and all(t["seconds"] > 0 for t in c["table"])
        assert autotune_cycle(amg, "fp32")["cached"]
        for w in (1, 2, 4):
            for u in (1, 2):
                amg.set_tune("fp32", w, u, True)
                z = amg.vcycle(r).double().cpu().numpy()
                assert _rel(z, z_ref) <= 1e-5, (w, u)
    finally:
        op.close()
        amg.close()
