"""Host side of the library without a GPU: the C-ABI loads and exports every
declared entry point, and the host parts of the path (mesh generators, node
graph, RCM, block conversions, smoothed-aggregation setup) match the golden
vectors, scipy and the reference's rules.  No device call is made here."""
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _golden_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


G = _golden_module()


def _declared():
    txt = open(os.path.join(ROOT, "include", "wnfea_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(wnfea_\w+)\s*\(", txt)))


# ---------------------------------------------------------------- C-ABI
def test_library_exports_every_declared_symbol():
    from wnfea_amd import _lib
    L = _lib.load()
    declared = _declared()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(L, name), name
    # the ctypes signature table covers exactly the declared entry points
    assert sorted(_lib.SIGNATURES) == declared
    assert L.wnfea_abi_version() == 1


def test_header_arity_matches_ctypes_table():
    from wnfea_amd import _lib
    txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "wnfea_hip.h")).read(), flags=re.S)
    for name, (_, args) in _lib.SIGNATURES.items():
        m = re.search(r"\b%s\s*\(([^)]*)\)" % name, txt)
        assert m, name
        params = [p for p in m.group(1).split(",") if p.strip() and p.strip() != "void"]
        assert len(params) == len(args), (name, len(params), len(args))


# ---------------------------------------------------------------- meshes
def test_c0_generator_matches_config():
    model = G.c0_model()
    assert len(model.mesh_nodes) == 6561 and len(model.solid_elements) == 3840
    assert model.solid_elements.dtype == np.int32 and model.solid_elements.shape[1] == 10
    X = model.mesh_nodes[model.solid_elements]
    # midside nodes at the edge midpoints (Abaqus order m01 m12 m20 m03 m13 m23)
    for m, (a, b) in enumerate(G.EDGES):
        assert np.abs(X[:, 4 + m] - 0.5 * (X[:, a] + X[:, b])).max() < 1e-15
    # every lattice point of the 81x9x9 half grid is a node exactly once
    q = np.round(model.mesh_nodes / np.array([1.0 / 80, 0.1 / 8, 0.1 / 8])).astype(int)
    assert len({tuple(r) for r in q}) == 6561


def test_ibeam_loads_resultants():
    """Tip shear splits equally; the torsion couple has zero resultant force
    and moment T about the section centroid."""
    from wnfea_amd.mesh.solid_generator import generate_c3d10_ibeam
    model, meta = generate_c3d10_ibeam(n_layers=4, length=0.04, shear_total=-50e3, torque=5e3)
    F = model.load_vector(3).reshape(-1, 3)
    tip = meta["tip_nodes"]
    assert abs(F[:, 2].sum() + 50e3) < 1e-6 and abs(F[:, 1].sum()) < 1e-6 and abs(F[:, 0].sum()) == 0.0
    ry = model.mesh_nodes[tip, 1]
    rz = model.mesh_nodes[tip, 2] - 0.15
    Mx = np.sum(ry * F[tip, 2] - rz * F[tip, 1])
    assert abs(Mx - 5e3) < 1e-6
    assert meta["section_quads"] == 116
    assert model.fixed_dof_mask(3).sum() == 3 * meta["root_nodes_fixed"]
    X = model.mesh_nodes[model.solid_elements]
    vol = np.einsum("mi,mi->m", np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), X[:, 3] - X[:, 0]) / 6
    assert np.all(vol > 0)


def test_ibeam_full_size_counts():
    """The 300-layer I-beam: 351 585 nodes, 208 800 tets, 1 054 755 DOF."""
    from wnfea_amd.mesh.solid_generator import generate_c3d10_ibeam
    _, meta = generate_c3d10_ibeam()
    assert (meta["n_nodes"], meta["n_elements"], meta["n_dofs"]) == (351585, 208800, 1054755)


# ---------------------------------------------------------------- graph and RCM
def test_node_graph_matches_reference_construction():
    from wnfea_amd.mesh.rcm import compute_node_adjacency_matrix
    model = G.c0_model()
    A = compute_node_adjacency_matrix(len(model.mesh_nodes), None, model.solid_elements)
    R = G.node_graph(len(model.mesh_nodes), model.solid_elements)
    R.sort_indices()
    assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
    assert A.dtype == np.int8 and np.all(A.data == 1)


def test_rcm_matches_golden_and_scipy():
    from wnfea_amd.mesh.rcm import compute_node_adjacency_matrix, compute_rcm_order
    model = G.c0_model()
    A = compute_node_adjacency_matrix(len(model.mesh_nodes), None, model.solid_elements)
    n2o, o2n = compute_rcm_order(A)
    assert np.array_equal(n2o, np.load(os.path.join(GOLDEN, "c0_rcm_new_to_old.npy")))
    assert np.array_equal(o2n[n2o], np.arange(len(n2o)))
    rng = np.random.default_rng(11)
    for t in range(40):
        n = int(rng.integers(1, 150))
        B = sp.random(n, n, density=float(rng.uniform(0.0, 0.1)), random_state=100 + t, format="csr")
        B = sp.csr_matrix(((B + B.T + sp.identity(n)) > 0).astype(np.int8))
        B.sort_indices()
        assert np.array_equal(compute_rcm_order(B)[0], csg.reverse_cuthill_mckee(B, symmetric_mode=True)), t


def test_apply_rcm_roundtrip_and_bandwidth():
    """tests/test_bsr_rcm.py:61-91: bandwidth does not grow; the permuted
    solution maps back exactly; supports and loads follow their nodes."""
    from oracle import oracle as O
    from wnfea_amd.mesh.rcm import apply_rcm_to_model, compute_matrix_bandwidth, permute_solution_to_original
    model, _ = G.generate_c3d10_structured_block(0.5, 0.1, 0.1, 20, 3, 3, tets_per_hex=6, load_direction="z")
    mr, n2o, o2n = apply_rcm_to_model(model)
    bw0 = compute_matrix_bandwidth(O.node_pattern(len(model.mesh_nodes), model.solid_elements))
    bw1 = compute_matrix_bandwidth(O.node_pattern(len(mr.mesh_nodes), mr.solid_elements))
    assert bw1 <= bw0
    for dpn in (3, 6):
        u = np.linspace(1.0, 100.0, len(n2o) * dpn)
        up = u[(n2o[:, None] * dpn + np.arange(dpn)).ravel()]
        assert np.array_equal(permute_solution_to_original(up, n2o, dpn), u)
    assert np.array_equal(mr.load_vector(3).reshape(-1, 3), model.load_vector(3).reshape(-1, 3)[n2o])
    assert np.array_equal(mr.fixed_dof_mask(3).reshape(-1, 3), model.fixed_dof_mask(3).reshape(-1, 3)[n2o])


# ---------------------------------------------------------------- block matrices
@pytest.mark.parametrize("b", [3, 6])
def test_csr_bsr_roundtrip_and_oracle_pattern(b):
    """tests/test_bsr_rcm.py:94-130 (round trip < 1e-15) and the block pattern
    of the oracle's restatement of bsr_matrix.py:49-122."""
    from oracle import oracle as O
    from wnfea_amd.solver.bsr_matrix import bsr_to_csr, csr_to_bsr
    model, _ = G.generate_c3d10_structured_block(0.3, 0.1, 0.1, 6, 2, 2, tets_per_hex=6, load_direction="z")
    E, nu = model.element_moduli()
    K, _ = O.assemble_csr(model.mesh_nodes, model.solid_elements, E, nu, model.fixed_dof_mask(3),
                          model.load_vector(3))
    n = K.shape[0]
    if n % b:
        K = sp.block_diag([K, sp.identity(b - n % b)]).tocsr()
    B = csr_to_bsr(K, b, dtype=np.float64)
    rp, ci, vals = O.csr_to_bsr(K, b)
    assert np.array_equal(np.asarray(B.block_row_ptr), rp) and np.array_equal(np.asarray(B.block_col_idx), ci)
    assert np.array_equal(np.asarray(B.values).reshape(len(ci), -1), vals)
    K2 = bsr_to_csr(B)
    assert sp.linalg.norm(K2 - K) / sp.linalg.norm(K) < 1e-15
    assert B.nnz_blocks == len(ci) and B.block_size == b and B.shape == K.shape


# ---------------------------------------------------------------- smoothed aggregation (host setup)
def test_sa_setup_invariants():
    """Aggregation covers every free node; R = P^T; coarse operators are the
    symmetric Galerkin products R A P; the coarsest is positive definite and
    its stored inverse is its inverse; P reproduces the coarse rigid-body
    modes on the tentative level."""
    from oracle import oracle as O
    from wnfea_amd.solver.amg_preconditioner import sa_setup
    model, _ = G.generate_c3d10_structured_block(1.0, 0.1, 0.1, 24, 3, 3, tets_per_hex=6, load_direction="z")
    E, nu = model.element_moduli()
    fixed = model.fixed_dof_mask(3)
    K, _ = O.assemble_csr(model.mesh_nodes, model.solid_elements, E, nu, fixed, model.load_vector(3))
    d = K.diagonal()
    s = 1.0 / np.sqrt(np.abs(d))
    A = (sp.diags(s) @ K @ sp.diags(s)).tobsr(blocksize=(3, 3))
    levels, inv = sa_setup(A, model.mesh_nodes, fixed, coarse_size=300, dof_scale=np.sqrt(np.abs(d)))
    assert len(levels) >= 2
    for lv, nxt in zip(levels[:-1], levels[1:]):
        assert nxt.A.shape[0] < lv.A.shape[0]
        assert abs(lv.R - lv.P.T).max() == 0.0
        Ac = (lv.R @ lv.A @ lv.P).toarray()
        assert np.abs(nxt.A.toarray() - 0.5 * (Ac + Ac.T)).max() <= 1e-12 * np.abs(Ac).max()
        # the Chebyshev bound covers the spectrum of D^-1 A, without a large margin
        dh = np.sqrt(lv.dinv)
        lam = sp.linalg.eigsh(sp.diags(dh) @ lv.A.tocsr() @ sp.diags(dh), k=1, which="LA",
                              return_eigenvectors=False)[0]
        assert lam <= lv.lmax <= 1.25 * lam
    Acd = levels[-1].A.toarray()
    assert np.linalg.eigvalsh(Acd).min() > 0
    assert np.abs(inv @ Acd - np.eye(len(Acd))).max() < 1e-8


def test_vcycle_mirror_pcg_converges():
    """The numpy restatement of the library's V-cycle (tests/vcycle_mirror.py)
    is a symmetric positive preconditioner that makes CG converge fast; the
    GPU test compares the device cycle with it."""
    from oracle import oracle as O
    from vcycle_mirror import pcg, vcycle
    from wnfea_amd.solver.amg_preconditioner import sa_setup
    model, _ = G.generate_c3d10_structured_block(1.0, 0.1, 0.1, 24, 3, 3, tets_per_hex=6, load_direction="z")
    E, nu = model.element_moduli()
    fixed = model.fixed_dof_mask(3)
    K, F = O.assemble_csr(model.mesh_nodes, model.solid_elements, E, nu, fixed, model.load_vector(3))
    d = np.abs(K.diagonal())
    s = 1.0 / np.sqrt(d)
    A = (sp.diags(s) @ K @ sp.diags(s)).tocsr()
    levels, inv = sa_setup(A.tobsr(blocksize=(3, 3)), model.mesh_nodes, fixed, coarse_size=300,
                           dof_scale=np.sqrt(d))
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    Mx, My = vcycle(levels, inv, x), vcycle(levels, inv, y)
    assert abs(x @ My - y @ Mx) <= 1e-10 * abs(x @ Mx) and x @ Mx > 0
    u_hat, it = pcg(A, s * F, lambda r: vcycle(levels, inv, r), 1e-8)
    u = s * u_hat
    assert it < 60
    assert np.linalg.norm(F - K @ u) / np.linalg.norm(F) < 1e-7


# ---------------------------------------------------------------- many-mesh batches
def test_box_batch_generator_and_merge():
    """Seeded uneven boxes inside the DOF window; the merged model is the
    block-diagonal union (offsets, loads, supports, moduli carried along)."""
    from wnfea_amd.solver.batch import generate_box_batch, merge_models
    batch = generate_box_batch(8, seed=1, dof_range=(2000, 60000))
    again = generate_box_batch(8, seed=1, dof_range=(2000, 60000))
    assert [b[1]["grid"] for b in batch] == [b[1]["grid"] for b in again]
    for m, meta in batch:
        nx, ny, nz = meta["grid"]
        assert 10 <= nx <= 60 and 3 <= ny <= 8 and ny == nz
        assert 2000 <= meta["n_dofs"] == 3 * (2 * nx + 1) * (2 * ny + 1) ** 2 <= 60000
    models = [b[0] for b in batch]
    merged, offs = merge_models(models)
    assert offs[-1] == len(merged.mesh_nodes) == sum(len(m.mesh_nodes) for m in models)
    F = merged.load_vector(3)
    fixed = merged.fixed_dof_mask(3)
    for k, m in enumerate(models):
        sl = slice(3 * offs[k], 3 * offs[k + 1])
        assert np.array_equal(F[sl], m.load_vector(3))
        assert np.array_equal(fixed[sl], m.fixed_dof_mask(3))
    conn = merged.solid_elements
    owner = np.searchsorted(offs, conn, side="right") - 1
    assert np.all(owner == owner[:, :1])  # no element spans two meshes
    E, nu = merged.element_moduli()
    assert len(E) == len(conn) and np.all(E == 2.1e11)


def test_ibeam_tip_load_cases_resultants():
    from wnfea_amd.mesh.solid_generator import generate_c3d10_ibeam, ibeam_tip_load_cases
    model, meta = generate_c3d10_ibeam(n_layers=2, length=0.02)
    tip = meta["tip_nodes"]
    B = ibeam_tip_load_cases(model.mesh_nodes, tip, n_cases=4, seed=0)
    rng = np.random.default_rng(0)
    X = model.mesh_nodes
    for k in range(4):
        F, M = rng.normal(0.0, 20e3, 3), rng.normal(0.0, 2e3, 3)
        f = B[:, k].reshape(-1, 3)
        assert np.allclose(f.sum(axis=0), F, rtol=0, atol=1e-6)
        r = X - np.array([0.0, 0.0, 0.15])
        r[:, 0] = 0.0  # moments about the tip-section centroid
        assert np.allclose(np.cross(r, f).sum(axis=0), M, rtol=0, atol=1e-6)


def test_precision_config_presets():
    from wnfea_amd.solver import mixed_precision as mp
    assert mp.MIXED_FP32.amg_precision == "fp32" and mp.FULL_FP64.amg_precision == "fp64"
    assert mp.TRI_PRECISION.use_tri_precision and mp.TRI_PRECISION.switch_tol == 1e-2
    assert "Tri-precision" in mp.TRI_PRECISION.summary() and "Uniform" in mp.FULL_FP64.summary()
    assert abs(mp.MIXED_FP32.inner_eps_scale - np.sqrt(np.finfo(np.float32).eps)) == 0.0


@pytest.mark.parametrize("tets_per_hex", [5, 6])
def test_sa_setup_rank_deficient_aggregates(tets_per_hex):
    """A high strength threshold leaves aggregates of one or two nodes that
    cannot carry all six rigid-body modes: P gets zero columns and the
    Galerkin product zero rows.  Setup must still give a nonsingular coarse
    level (unit diagonal on the vanished rows, as amg_preconditioner.py:288-293
    of the reference) whose stored inverse is its inverse, and the cycle must
    still precondition CG."""
    from oracle import oracle as O
    from vcycle_mirror import pcg, vcycle
    from wnfea_amd.solver.amg_preconditioner import sa_setup
    model, _ = G.generate_c3d10_structured_block(1.0, 0.1, 0.1, 40, 4, 4, tip_load_total=-1000.0,
                                                 tets_per_hex=tets_per_hex, load_direction="z")
    E, nu = model.element_moduli()
    fixed = model.fixed_dof_mask(3)
    K, F = O.assemble_csr(model.mesh_nodes, model.solid_elements, E, nu, fixed, model.load_vector(3))
    d = np.abs(K.diagonal())
    s = 1.0 / np.sqrt(d)
    A = (sp.diags(s) @ K @ sp.diags(s)).tocsr()
    levels, inv = sa_setup(A.tobsr(blocksize=(3, 3)), model.mesh_nodes, fixed, theta=0.25, coarse_size=1000,
                           dof_scale=np.sqrt(d))
    assert len(levels) >= 2
    P = levels[0].P.tocsc()
    colnorm = np.sqrt(np.asarray(P.multiply(P).sum(axis=0))).ravel()
    vanished = np.flatnonzero(colnorm == 0.0)
    assert len(vanished) > 0  # the case this test is about
    A1 = levels[1].A.tocsr()
    assert np.all(A1.diagonal()[vanished] == 1.0)
    Acd = levels[-1].A.toarray()
    assert np.linalg.eigvalsh(Acd).min() > 0
    assert np.abs(inv @ Acd - np.eye(len(Acd))).max() < 1e-8
    u_hat, it = pcg(A, s * F, lambda r: vcycle(levels, inv, r), 1e-8)
    assert it < 400  # a poor hierarchy (threshold 0.25), but a convergent one
    assert np.linalg.norm(F - K @ (s * u_hat)) / np.linalg.norm(F) < 1e-7


def test_reference_import_path():
    """Reference callers' imports (tests/test_solid_bench.py:12-13,
    tests/test_bsr_rcm.py:24-35, test_c3d10.py:11-19, wnfea/__init__.py:7-31)
    resolve unchanged to this implementation's modules, one module object per
    name; modules off the solve path stay absent."""
    import importlib
    import wnfea
    from wnfea.solver.solid_solver import (assemble_c3d10_sparse_3dof, compute_c3d10_stress_field,  # noqa: F401
                                           solve_c3d10_pcg, solve_c3d10_system)
    from wnfea.mesh.rcm import (apply_rcm_to_model, compute_matrix_bandwidth,  # noqa: F401
                                compute_node_adjacency_matrix, compute_rcm_order, permute_solution_to_original)
    from wnfea.solver.bsr_matrix import bsr6x6_to_csr, csr_to_bsr6x6, hip_spmv_bsr6x6  # noqa: F401
    from wnfea.elements.c3d10 import (b_matrix_c3d10, element_internal_forces_c3d10,  # noqa: F401
                                      element_stiffness_c3d10, element_stresses_c3d10, jacobian_c3d10,
                                      shape_functions_c3d10)
    from wnfea.mesh.solid_generator import generate_c3d10_structured_block  # noqa: F401
    from wnfea.solver.fast_kernels import compute_internal_forces_fast  # noqa: F401
    from wnfea.solver.assembler import assemble_global_system_sparse  # noqa: F401
    from wnfea.solver.residual import build_external_force_vector, compute_equilibrium_residual  # noqa: F401
    from wnfea.solver.nonlinear_solver import pcg_solve, solve_nonlinear_jfnk  # noqa: F401
    from wnfea.solver.dof_manager import DOFManager  # noqa: F401
    from wnfea.boundary.conditions import LoadDef, create_fixed_support  # noqa: F401
    from wnfea.model import FEAModel, PropertyAssignment  # noqa: F401
    from wnfea import BlockBeamAMGPreconditioner, FEAModel as F2, MIXED_FP32  # noqa: F401
    for name in ("solver.solid_solver", "mesh.rcm", "elements.c3d10", "model", "solver.amg_preconditioner"):
        assert importlib.import_module("wnfea." + name) is importlib.import_module("wnfea_amd." + name)
    assert wnfea.solver.solid_solver.solve_c3d10_system is solve_c3d10_system and F2 is FEAModel
    with pytest.raises(ImportError):
        importlib.import_module("wnfea.geometry.step_parser")


@pytest.mark.parametrize("tets_per_hex", [5, 6])
def test_structured_block_conformity(tets_per_hex):
    """The reference-default 5-tet split (solid_generator.py:66-163) and the
    6-tet split: node count of the half lattice actually used (5 921 for the
    5-tet 40x4x4 box: the 5-tet split uses only some face diagonals), every
    interior triangle shared by exactly two tets with opposite orientation,
    every boundary triangle on the box surface, positive volumes, total
    volume L*W*H, midside nodes at edge midpoints and shared by every tet on
    the edge."""
    from wnfea_amd.mesh.solid_generator import generate_c3d10_structured_block
    L, W, H = 1.0, 0.1, 0.1
    model, meta = generate_c3d10_structured_block(L, W, H, 40, 4, 4, tip_load_total=-1000.0,
                                                  tets_per_hex=tets_per_hex, load_direction="z")
    X, C = model.mesh_nodes, model.solid_elements
    assert len(C) == tets_per_hex * 40 * 4 * 4
    assert len(X) == (5921 if tets_per_hex == 5 else 6561)
    assert len(np.unique(C)) == len(X)  # every node is used
    P = X[C]
    vol = np.einsum("mi,mi->m", np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), P[:, 3] - P[:, 0]) / 6.0
    assert vol.min() > 0.0
    assert abs(vol.sum() - L * W * H) < 1e-12
    for m, (a, b) in enumerate(G.EDGES):
        assert np.abs(P[:, 4 + m] - 0.5 * (P[:, a] + P[:, b])).max() < 1e-15
    # edge -> midside node is a function of the edge (conforming quadratic mesh)
    ends = np.sort(np.stack([C[:, [a, b]] for a, b in G.EDGES], axis=1).reshape(-1, 2), axis=1)
    mids = C[:, 4:10].reshape(-1)
    key = ends[:, 0].astype(np.int64) * len(X) + ends[:, 1]
    order = np.argsort(key, kind="stable")
    k, md = key[order], mids[order]
    same = k[1:] == k[:-1]
    assert np.all(md[1:][same] == md[:-1][same])
    # faces: oriented outward from each tet (vertex order 0-1-2-3 has positive volume)
    faces = np.concatenate([C[:, [0, 2, 1]], C[:, [0, 1, 3]], C[:, [1, 2, 3]], C[:, [0, 3, 2]]])
    srt = np.sort(faces, axis=1)
    _, inv, cnt = np.unique(srt, axis=0, return_inverse=True, return_counts=True)
    inv = inv.ravel()
    assert cnt.max() == 2
    # interior faces appear twice with opposite orientation (an even and an odd
    # permutation of the sorted triple)
    def parity(f):
        s = np.argsort(f, axis=1)
        inv_count = (s[:, 0] > s[:, 1]).astype(int) + (s[:, 0] > s[:, 2]) + (s[:, 1] > s[:, 2])
        return inv_count % 2
    par = parity(faces)
    twice = np.flatnonzero(cnt == 2)
    first = np.full(len(cnt), -1)
    second = np.full(len(cnt), -1)
    for i, f in enumerate(inv):
        if first[f] < 0:
            first[f] = i
        else:
            second[f] = i
    assert np.all(par[first[twice]] != par[second[twice]])
    # boundary faces lie on the box surface
    once = first[cnt == 1]
    c = X[faces[once]]  # (nb, 3, 3)
    on = np.zeros(len(once), dtype=bool)
    for ax, ext in enumerate((L, W, H)):
        for v in (0.0, ext):
            on |= np.all(np.abs(c[:, :, ax] - v) < 1e-12, axis=1)
    assert on.all()
    boundary_area = 2 * (L * W + L * H + W * H)
    area = 0.5 * np.linalg.norm(np.cross(c[:, 1] - c[:, 0], c[:, 2] - c[:, 0]), axis=1).sum()
    assert abs(area - boundary_area) < 1e-12


def test_coarse_inverse_rejects_an_empty_level():
    """The coarsest-level set-up refuses n < 1 (the packed tile product reads
    the vector's last entry for its padding lanes) before touching the GPU."""
    import ctypes
    from wnfea_amd import _lib
    L = _lib.load()
    h = L.wnfea_amg_create(2, 3)
    assert h
    try:
        inv = (ctypes.c_double * 1)(1.0)
        assert L.wnfea_amg_set_coarse(h, 0, inv) == -1
        assert L.wnfea_amg_set_coarse(h, 1, None) == -1
    finally:
        L.wnfea_amg_destroy(h)


def test_equilibrated_finest_level_has_a_unit_fp32_inverse_diagonal():
    """On S K S with S = diag(K)^-1/2 every diagonal entry is K_ii / K_ii:
    the finest level's point-Jacobi inverse rounds to exactly 1.0 in FP32
    (the matrix-free Chebyshev steps then skip its load), and a level below
    it does not."""
    from oracle import oracle as O
    from wnfea_amd.solver.amg_preconditioner import sa_setup
    model, _ = G.generate_c3d10_structured_block(1.0, 0.1, 0.1, 24, 3, 3, tets_per_hex=6, load_direction="z")
    E, nu = model.element_moduli()
    fixed = model.fixed_dof_mask(3)
    K, _ = O.assemble_csr(model.mesh_nodes, model.solid_elements, E, nu, fixed, model.load_vector(3))
    d = K.diagonal()
    s = 1.0 / np.sqrt(np.abs(d))
    A = (sp.diags(s) @ K @ sp.diags(s)).tobsr(blocksize=(3, 3))
    levels, _ = sa_setup(A, model.mesh_nodes, fixed, coarse_size=300, dof_scale=np.sqrt(np.abs(d)))
    assert np.all(levels[0].dinv.astype(np.float32) == np.float32(1.0))
    assert np.abs(levels[0].dinv - 1.0).max() < 1e-15
    assert not np.all(levels[1].dinv.astype(np.float32) == np.float32(1.0))
