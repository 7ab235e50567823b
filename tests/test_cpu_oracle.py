"""Pins of the CPU oracle (no GPU): golden vectors made by independent code
(tests/golden/make_golden.py), the reference's own compiled f_int, scipy's
RCM, closed forms and the invariants of the reference's test_c3d10.py."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg
from scipy.sparse.linalg import spsolve

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


G = _golden_module()


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def reference_tet(scale=1.0):
    c = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float) * scale
    mids = [0.5 * (c[a] + c[b]) for a, b in G.EDGES]
    return np.vstack([c, mids])


@pytest.fixture(scope="module")
def c0():
    model = G.c0_model()
    E, nu = model.element_moduli()
    K, F = O.assemble_csr(model.mesh_nodes, model.solid_elements, E, nu, model.fixed_dof_mask(3),
                          model.load_vector(3))
    return model, K, F


# ---------------------------------------------------------------- element
def test_shape_functions_kronecker_and_unity():
    """test_c3d10.py:46-87: N_i(node_j) = delta_ij, sum N = 1, sum dN = 0; dN
    equals an independent volume-coordinate derivation."""
    nat = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (.5, 0, 0), (.5, .5, 0), (0, .5, 0), (0, 0, .5), (.5, 0, .5),
           (0, .5, .5)]
    for j, p in enumerate(nat):
        N, _ = O.shape_derivs(*p)
        assert np.abs(N - np.eye(10)[j]).max() < 1e-14
    rng = np.random.default_rng(42)
    for _ in range(20):
        L = rng.dirichlet(np.ones(4))
        N, dN = O.shape_derivs(L[1], L[2], L[3])
        assert abs(N.sum() - 1.0) < 1e-14
        assert np.abs(dN.sum(axis=1)).max() < 1e-13
        assert np.abs(dN - G.dshape(L[1], L[2], L[3])).max() < 1e-14


def test_ke_matches_collapsed_gauss_rule():
    """Straight-edged tets: the 4-point Hammer rule is exact for B^T D B, so Ke
    equals the 125-point collapsed-cube Gauss rule to round-off (golden)."""
    g = np.load(os.path.join(GOLDEN, "element_sample.npz"))
    for X, Kg in zip(g["X_straight"], g["Ke_straight"]):
        Ke = O.element_stiffness(X, float(g["E_straight"]), float(g["nu_straight"]))
        assert np.abs(Ke - Kg).max() <= 1e-12 * np.abs(Kg).max()


def test_ke_rigid_modes_and_spectrum():
    """test_c3d10.py:90-143: symmetric, 6 zero and 24 positive eigenvalues,
    translations and rotations give zero force (straight and curved tets)."""
    X = reference_tet(1.5)
    Xc = X.copy()
    Xc[4] += [0.02, 0.01, -0.01]
    Xc[7] += [-0.01, 0.02, 0.01]
    for coords in (X, Xc):
        Ke = O.element_stiffness(coords, 2.1e11, 0.3)
        assert np.abs(Ke - Ke.T).max() <= 1e-15 * np.abs(Ke).max()
        ev = np.linalg.eigvalsh(Ke)
        tr = np.trace(Ke)
        assert np.all(np.abs(ev[:6]) / tr < 1e-12) and np.all(ev[6:] > 1e-8 * tr)
        for d in range(3):
            u = np.zeros(30)
            u[d::3] = 1.0
            assert np.linalg.norm(Ke @ u) / tr < 1e-12
        for axis in range(3):
            w = np.zeros(3)
            w[axis] = 1.0
            u = np.cross(w, coords).ravel()
            assert np.linalg.norm(Ke @ u) / tr < 1e-12


def test_constant_strain_patch():
    """test_c3d10.py:146-194: a linear field gives the exact constant stress."""
    X = reference_tet(2.0)
    X[4] += [0.02, 0.01, -0.01]
    X[7] += [-0.01, 0.02, 0.01]
    E, nu = 2.0e11, 0.28
    eps = np.array([1.2e-3, -0.5e-3, 0.8e-3, 0.6e-3, -0.4e-3, 0.3e-3])
    sig = G.dmat(E, nu) @ eps
    grad = np.array([[eps[0], 0.5 * eps[3], 0.5 * eps[5]], [0.5 * eps[3], eps[1], 0.5 * eps[4]],
                     [0.5 * eps[5], 0.5 * eps[4], eps[2]]])
    u = (X @ grad.T).ravel()
    s, vm, vmax = O.element_stresses(X, u, E, nu)
    assert np.abs(s - sig).max() <= 1e-10 * np.abs(sig).max()
    vm_ref = np.sqrt(0.5 * ((sig[0] - sig[1]) ** 2 + (sig[1] - sig[2]) ** 2 + (sig[2] - sig[0]) ** 2
                            + 6 * (sig[3] ** 2 + sig[4] ** 2 + sig[5] ** 2)))
    assert np.abs(vm - vm_ref).max() <= 1e-10 * vm_ref and abs(vmax - vm.max()) == 0.0


def test_fint_equals_ke_u():
    """test_c3d10.py:197-217: matrix-free f_int = Ke u_e to 1e-14."""
    X = reference_tet(1.2)
    Ke = O.element_stiffness(X, 1.8e11, 0.33)
    u = np.random.default_rng(99).standard_normal(30) * 1e-3
    assert _rel(O.element_internal_forces(X, u, 1.8e11, 0.33), Ke @ u) < 1e-14


def test_fint_matches_reference_kernels_golden():
    """Curved tets: oracle f_int against the reference's compiled kernels.cpp
    (vectors committed by make_golden.py)."""
    g = np.load(os.path.join(GOLDEN, "element_sample.npz"))
    for X, u, f in zip(g["X_curved"], g["u_curved"], g["f_curved_ref"]):
        assert _rel(O.element_internal_forces(X, u, float(g["E_curved"]), float(g["nu_curved"])), f) < 1e-13


def test_fint_global_matches_reference_kernels_live():
    """Stride-6 global f_int on C0-sized data against oracle/_ref (built from
    the reference source when the reference tree is present)."""
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    model, _ = G.generate_c3d10_structured_block(0.3, 0.1, 0.1, 6, 2, 2, tets_per_hex=6)
    nodes, conn = model.mesh_nodes, model.solid_elements.astype(np.int32)
    rng = np.random.default_rng(5)
    props = np.column_stack([rng.uniform(1e11, 2e11, len(conn)), rng.uniform(0.2, 0.35, len(conn))])
    U = rng.standard_normal(len(nodes) * 6) * 1e-4
    F = O.fint_global(nodes, conn, props, U, stride=6)
    Fr = np.zeros_like(F)
    p = np.ascontiguousarray(props)
    R.compute_c3d10_internal_forces_native(nodes.ctypes.data, conn.ctypes.data, p.ctypes.data, U.ctypes.data,
                                           Fr.ctypes.data, len(conn), len(nodes))
    assert _rel(F, Fr) < 1e-13


# ---------------------------------------------------------------- mesh, assembly, BCs
def test_c0_mesh_counts(c0):
    model, _, F = c0
    meta = json.load(open(os.path.join(GOLDEN, "golden.json")))["c0"]
    assert len(model.mesh_nodes) == 81 * 9 * 9 == meta["n_nodes"]
    assert len(model.solid_elements) == 40 * 4 * 4 * 6 == meta["n_elements"]
    assert model.fixed_dof_mask(3).sum() == 243 == meta["n_fixed_dofs"]
    assert abs(model.load_vector(3)[2::3].sum() + 1000.0) < 1e-9
    X = model.mesh_nodes[model.solid_elements]
    vol = np.einsum("mi,mi->m", np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), X[:, 3] - X[:, 0]) / 6
    assert np.all(vol > 0) and abs(vol.sum() - 1.0 * 0.1 * 0.1) < 1e-12


def test_c0_assembly_pattern_and_solution_golden(c0):
    """Oracle CSR (after BCs and eliminate_zeros) has the golden 3x3 block
    pattern (the block rule of bsr_matrix.py:84-86 applied to it gives the same
    blocks) and significant entries, and its exact solution is the golden
    (independently assembled) solution."""
    _, K, F = c0
    meta = json.load(open(os.path.join(GOLDEN, "golden.json")))["c0"]
    K = sp.csr_matrix(K)
    K.sort_indices()
    assert G.significant_nnz(K) == meta["csr_significant_nnz"]
    assert G.pattern_hash(G.block_pattern(K)) == meta["bsr3_pattern_sha256"]
    rp, ci, _ = O.csr_to_bsr(K, 3)
    Pb = sp.csr_matrix((np.ones(len(ci), dtype=np.int8), ci, rp), shape=(K.shape[0] // 3,) * 2)
    assert G.pattern_hash(Pb) == meta["bsr3_pattern_sha256"] and Pb.nnz == meta["bsr3_nnzb"]
    u = spsolve(K.tocsc(), F)
    ug = np.load(os.path.join(GOLDEN, "c0_solution.npz"))["u"]
    # two exact solves of matrices that differ by quadrature round-off (~1e-16
    # relative), amplified by the condition number of K
    assert _rel(u, ug) < 1e-9


def test_tiny_assembly_vs_dense_brute_force():
    """2x1x1 box: oracle CSR equals a dense element-by-element assembly with
    the BC rule applied entry by entry (solid_solver.py:95-127)."""
    model, _ = G.generate_c3d10_structured_block(0.2, 0.1, 0.1, 2, 1, 1, tets_per_hex=6, load_direction="z")
    n = len(model.mesh_nodes)
    E, nu = model.element_moduli()
    fixed = model.fixed_dof_mask(3)
    K, F = O.assemble_csr(model.mesh_nodes, model.solid_elements, E, nu, fixed, model.load_vector(3))
    Kd = np.zeros((3 * n, 3 * n))
    for e, conn in enumerate(model.solid_elements):
        Ke = G.ke_collapsed(model.mesh_nodes[conn][None], E[e], nu[e])[0]
        d = (3 * conn[:, None] + np.arange(3)).ravel()
        Kd[np.ix_(d, d)] += Ke
    Kd[fixed, :] = 0.0
    Kd[:, fixed] = 0.0
    Kd[fixed, fixed] = 1.0
    assert np.abs(K.toarray() - Kd).max() <= 1e-12 * np.abs(Kd).max()
    # the pattern holds exactly the (significant) non-zeros that survive the BC rule
    tol = 1e-9 * np.abs(Kd).max()
    assert G.significant_nnz(K) == np.count_nonzero(np.abs(Kd) > tol)
    Fd = model.load_vector(3)
    Fd[fixed] = 0.0
    assert np.array_equal(F, Fd)


def test_csr_to_bsr_pattern_rule():
    """Block pattern = union of block columns of the b rows, plus a diagonal
    block for an empty block row (bsr_matrix.py:84-86); values round-trip."""
    rng = np.random.default_rng(0)
    A = sp.random(60, 60, density=0.05, random_state=1, format="csr")
    A = A + A.T
    A[30:36, :] = 0.0  # an empty block row for b = 6
    A.eliminate_zeros()
    for b in (3, 6):
        rp, ci, vals = O.csr_to_bsr(A, b)
        dense_blocks = np.abs(A.toarray()).reshape(60 // b, b, 60 // b, b).sum(axis=(1, 3)) > 0
        for i in range(60 // b):
            cols = set(ci[rp[i]:rp[i + 1]])
            expect = set(np.nonzero(dense_blocks[i])[0]) or {i}
            assert cols == expect
        M = sp.bsr_matrix((vals.reshape(-1, b, b), ci, rp), shape=A.shape)
        assert np.abs(M.toarray() - A.toarray()).max() == 0.0
        x = rng.standard_normal(60)
        assert _rel(O.bsr_spmv(rp, ci, vals, x, b), A @ x) < 1e-14


def test_csr_spmv_vs_scipy(c0):
    _, K, _ = c0
    x = np.random.default_rng(1).standard_normal(K.shape[0])
    assert _rel(O.csr_spmv(K, x), K @ x) < 1e-14


# ---------------------------------------------------------------- solvers, stresses
def test_oracle_pcg_and_ir_vs_golden(c0):
    """FP64 Jacobi-PCG and the FP32-inner refinement loop reach the golden
    solution (tests/test_solid_bench.py:37-44 compares PCG with spsolve)."""
    _, K, F = c0
    ug = np.load(os.path.join(GOLDEN, "c0_solution.npz"))["u"]
    x, it, rel = O.pcg(K, F, rtol=1e-10, max_iter=20000)
    # rel is the true residual; the recursive one met 1e-10 (it < max_iter), the
    # true one sits near the FP64 floor of this system (spsolve: ~3e-10)
    assert it < 20000 and rel <= 1e-8
    assert abs(rel - np.linalg.norm(F - K @ x) / np.linalg.norm(F)) <= 1e-12
    assert _rel(x, ug) <= 1e-7
    x2, it2, rel2 = O.ir_mixed(K, F, rtol=1e-8, max_iter=20000)
    assert np.linalg.norm(F - K @ x2) / np.linalg.norm(F) <= 1e-8 and _rel(x2, ug) <= 1e-6


def test_stress_field_vs_golden(c0):
    model, _, _ = c0
    g = np.load(os.path.join(GOLDEN, "c0_solution.npz"))
    E, nu = model.element_moduli()
    cs, cv, nv = O.stress_field(model.mesh_nodes, model.solid_elements, E, nu, g["u"])
    assert np.abs(cs - g["cell_sig"]).max() <= 1e-9 * np.abs(g["cell_sig"]).max()
    assert np.abs(cv - g["cell_vm"]).max() <= 1e-9 * g["cell_vm"].max()
    assert np.abs(nv - g["nodal_vm"]).max() <= 1e-9 * g["nodal_vm"].max()


# ---------------------------------------------------------------- RCM
def test_oracle_rcm_matches_golden_and_scipy():
    """C0 node graph: oracle RCM = the committed scipy permutation; random
    symmetric graphs (with isolated nodes and several components): oracle =
    scipy.sparse.csgraph.reverse_cuthill_mckee."""
    model = G.c0_model()
    adj = G.node_graph(len(model.mesh_nodes), model.solid_elements)
    golden = np.load(os.path.join(GOLDEN, "c0_rcm_new_to_old.npy"))
    assert np.array_equal(O.rcm(adj), golden)
    rng = np.random.default_rng(7)
    for t in range(60):
        n = int(rng.integers(1, 120))
        A = sp.random(n, n, density=float(rng.uniform(0.0, 0.15)), random_state=t, format="csr")
        A = ((A + A.T) > 0).astype(np.int8)
        if t % 2:
            A = (A + sp.identity(n, dtype=np.int8, format="csr")).astype(np.int8)
        A = sp.csr_matrix(A)
        A.sort_indices()
        ref = csg.reverse_cuthill_mckee(A, symmetric_mode=True)
        assert np.array_equal(O.rcm(A), ref), t


# ---------------------------------------------------------------- sensitivities
def test_adjoint_gradient_matches_central_differences():
    """dC/dE_e (compliance) and d(tip u_z)/dE_e from the closed form against
    central differences (h = 1e-4 E_e, FP64 exact solves) on a small box with
    seeded random moduli."""
    model, _ = G.generate_c3d10_structured_block(0.4, 0.1, 0.1, 8, 2, 2, tip_load_total=-1000.0, tets_per_hex=6,
                                                 load_direction="z")
    M = len(model.solid_elements)
    rng = np.random.default_rng(2)
    E = 210e9 * (0.2 + 0.8 * rng.uniform(size=M))
    nu = np.full(M, 0.3)
    fixed = model.fixed_dof_mask(3)
    F = model.load_vector(3)
    tip = np.flatnonzero(np.abs(model.mesh_nodes[:, 0] - 0.4) < 1e-12)
    c = np.zeros(3 * len(model.mesh_nodes))
    c[3 * tip + 2] = 1.0 / len(tip)

    def solve(Ev, rhs):
        K, Fb = O.assemble_csr(model.mesh_nodes, model.solid_elements, Ev, nu, fixed, rhs)
        return spsolve(K.tocsc(), Fb)

    u = solve(E, F)
    lam = solve(E, c)
    gC = O.adjoint_grad(model.mesh_nodes, model.solid_elements, E, nu, u, u)
    gT = O.adjoint_grad(model.mesh_nodes, model.solid_elements, E, nu, u, lam)
    for e in rng.choice(M, 8, replace=False):
        h = 1e-4 * E[e]
        Ep, Em = E.copy(), E.copy()
        Ep[e] += h
        Em[e] -= h
        up, um = solve(Ep, F), solve(Em, F)
        fdC = (F @ up - F @ um) / (2 * h)
        fdT = (c @ up - c @ um) / (2 * h)
        assert abs(fdC - gC[e]) <= 1e-5 * abs(gC[e]), e
        assert abs(fdT - gT[e]) <= 1e-5 * np.abs(gT).max(), e
    assert np.all(gC <= 0)  # stiffer elements never increase compliance
