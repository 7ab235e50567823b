"""The batched (many tip load cases) and adjoint workloads on the I-beam
section, on shortened beams that a sparse direct factorization covers.

* 32 seeded tip load cases (resultant forces N(0, 20 kN) per component,
  moments N(0, 2 kN m) per component, seed 0) on a 30-layer I-beam
  (107 055 DOF) through the multi-column solver: every column meets 1e-8 on
  its FP64 residual and matches scipy's sparse LU solution of the oracle's K
  to 1e-6.
* Per-element moduli E_e = 210 GPa (0.2 + 0.8 U(0, 1)) (seed 2) on a
  20-layer I-beam (71 955 DOF): compliance and mean tip z-displacement and
  their gradients against the CPU FP64 reference gradient (oracle element
  matrices, exact solutions by sparse LU) to 1e-5, and the compliance
  gradient against central differences (h = 1e-4 E_e, FP64 forward solves)
  on 20 seeded elements (seed 3) to 1e-4 each.
* The V-cycle of a 3-level hierarchy (30 layers; a 6x6-block level 1 with
  its own smoothing) against the FP64 numpy mirror of the same hierarchy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def _short_ibeam(layers):
    from wnfea_amd.mesh.rcm import apply_rcm_to_model
    from wnfea_amd.mesh.solid_generator import generate_c3d10_ibeam
    model, meta = generate_c3d10_ibeam(n_layers=layers, length=3.0 * layers / 300)
    mr, n2o, o2n = apply_rcm_to_model(model)
    return mr, meta, o2n


@pytest.mark.timeout(300)
def test_32_tip_load_cases_match_direct_solutions():
    import torch
    from scipy.sparse.linalg import splu
    from oracle import oracle as O
    from wnfea_amd.mesh.solid_generator import ibeam_tip_load_cases
    from wnfea_amd.solver.solid_solver import assemble_c3d10_bsr, build_amg, solve_multi
    mr, meta, o2n = _short_ibeam(30)
    assert meta["n_dofs"] == 107055
    B = ibeam_tip_load_cases(mr.mesh_nodes, o2n[meta["tip_nodes"]], n_cases=32, seed=0)
    system = assemble_c3d10_bsr(mr)
    B[system.fixed_mask] = 0.0
    amg = build_amg(system, storage=("fp32",), theta=0.1, theta_decay=0.4)
    try:
        U, iters, rels, _ = solve_multi(system, amg, torch.as_tensor(B, device="cuda"), tol=1e-8, group=8)
    finally:
        amg.close()
    assert np.all(rels <= 1e-8) and np.all(iters > 0)
    E, nu = mr.element_moduli()
    K, _ = O.assemble_csr(mr.mesh_nodes, mr.solid_elements, E, nu, mr.fixed_dof_mask(3), mr.load_vector(3))
    lu = splu(K.tocsc())
    X = lu.solve(B)
    Uh = U.cpu().numpy()
    for k in range(32):
        assert np.linalg.norm(B[:, k] - K @ Uh[:, k]) <= 1e-8 * np.linalg.norm(B[:, k]), k
        assert _rel(Uh[:, k], X[:, k]) <= 1e-6, k


@pytest.fixture(scope="module")
def c3_short():
    from scipy.sparse.linalg import splu
    from oracle import oracle as O
    from wnfea_amd.solver.adjoint import C3D10Adjoint, tip_vector
    mr, meta, o2n = _short_ibeam(20)
    M = len(mr.solid_elements)
    E = 210e9 * (0.2 + 0.8 * np.random.default_rng(2).uniform(size=M))
    nu = np.full(M, 0.3)
    mr.solid_E, mr.solid_nu = E, nu
    tip = o2n[meta["tip_nodes"]]
    c = tip_vector(mr, tip)
    K, F = O.assemble_csr(mr.mesh_nodes, mr.solid_elements, E, nu, mr.fixed_dof_mask(3), mr.load_vector(3))
    lu = splu(K.tocsc())
    u, lam = lu.solve(F), lu.solve(c)
    ref = {"gC": O.adjoint_grad(mr.mesh_nodes, mr.solid_elements, E, nu, u, u),
           "gT": O.adjoint_grad(mr.mesh_nodes, mr.solid_elements, E, nu, u, lam),
           "C": F @ u, "T": c @ u}
    prob = C3D10Adjoint(mr, tol=1e-11, amg_precision="fp32", amg_kw={"theta": 0.1, "theta_decay": 0.4})
    yield mr, E, tip, ref, prob
    prob.close()


@pytest.mark.timeout(300)
def test_adjoint_gradients_vs_cpu_fp64_reference(c3_short):
    import torch
    mr, E, tip, ref, prob = c3_short
    Et = torch.tensor(E, dtype=torch.float64, device="cuda", requires_grad=True)
    C = prob.compliance(Et)
    C.backward()
    assert abs(float(C.detach()) - ref["C"]) <= 1e-8 * abs(ref["C"])
    assert _rel(Et.grad.cpu().numpy(), ref["gC"]) <= 1e-5
    Et.grad = None
    T = prob.tip_displacement(Et, tip)
    T.backward()
    assert abs(float(T.detach()) - ref["T"]) <= 1e-8 * abs(ref["T"])
    assert _rel(Et.grad.cpu().numpy(), ref["gT"]) <= 1e-5
    # under no_grad the forward pays the forward solve only
    solves = prob.solves
    with torch.no_grad():
        prob.tip_displacement(Et, tip)
    assert prob.solves == solves + 1


@pytest.mark.timeout(300)
def test_compliance_gradient_vs_central_differences_20_elements(c3_short):
    mr, E, tip, ref, prob = c3_short
    _, g, _ = prob.compliance_and_grad(E)
    g = g.cpu().numpy()
    rng = np.random.default_rng(3)
    for e in rng.choice(len(E), 20, replace=False):
        h = 1e-4 * E[e]
        fd = prob.compliance_central_difference(E, int(e), h)
        assert abs(fd - g[e]) <= 1e-4 * abs(g[e]), (int(e), fd, g[e])


@pytest.mark.timeout(300)
def test_three_level_vcycle_matches_numpy_mirror():
    """The V-cycle on a 3-level hierarchy of the 30-layer I-beam (107 055 ->
    13 080 -> 1 008 rows: a 6x6-block level 1 with its own Chebyshev
    smoothing, defect and 6x6 transfers) equals the FP64 numpy mirror of the
    same hierarchy (tests/vcycle_mirror.py) to 1e-12 in FP64, and to 1e-10
    with the finest level's smoothing and defect on the matrix-free operator
    (equal to the assembled one to FP64 rounding); the FP32 cycle agrees to
    FP32 accuracy."""
    import torch
    from vcycle_mirror import vcycle
    from wnfea_amd.solver.solid_solver import MatrixFreeOperator, assemble_c3d10_bsr, build_amg
    mr, meta, o2n = _short_ibeam(30)
    system = assemble_c3d10_bsr(mr, want_fp64_equilibrated=True)
    amg = build_amg(system, storage=("fp64", "fp32"), theta=0.1, theta_decay=0.4)
    op = MatrixFreeOperator(system)
    try:
        assert [lv.A.shape[0] for lv in amg.levels] == [107055, 13080, 1008]
        assert amg.levels[1].A.blocksize == (6, 6)
        n = amg.levels[0].A.shape[0]
        r = np.random.default_rng(4).standard_normal(n)
        z_ref = vcycle(amg.levels, amg.coarse_inv, r, deg=amg.level_degrees, ratio=amg.ratio)
        z = amg.vcycle(torch.as_tensor(r, device="cuda")).cpu().numpy()
        assert _rel(z, z_ref) <= 1e-12
        z32 = amg.vcycle(torch.as_tensor(r, device="cuda").float()).double().cpu().numpy()
        assert _rel(z32, z_ref) <= 1e-4
        # the FP32 cycle with its coarsest solve on the FP64 inverse
        amg.set_coarse_precision("fp64")
        z32c = amg.vcycle(torch.as_tensor(r, device="cuda").float()).double().cpu().numpy()
        amg.set_coarse_precision("fp32")
        assert _rel(z32c, z_ref) <= 1e-4 and not np.array_equal(z32c, z32)
        op.attach_smoother(amg)
        zmf = amg.vcycle(torch.as_tensor(r, device="cuda")).cpu().numpy()
        assert _rel(zmf, z_ref) <= 1e-10
    finally:
        op.close()
        amg.close()
