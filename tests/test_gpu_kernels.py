"""Element, assembly and SpMV kernels on the GPU against the CPU FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _random_tets(n, seed=0):
    rng = np.random.default_rng(seed)
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    out = []
    while len(out) < n:
        A = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
        if np.linalg.det(A) <= 0.2:
            continue
        c = base @ A.T * rng.uniform(0.01, 2.0) + rng.standard_normal(3)
        mids = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
        X = np.vstack([c] + [0.5 * (c[a] + c[b]) + 0.02 * rng.standard_normal(3) * 0.1 for a, b in mids])
        out.append(X)
    return np.array(out)


def test_ke_matches_oracle():
    from oracle import oracle as O
    from wnfea_amd.elements.c3d10 import element_stiffness_c3d10
    X = _random_tets(300)
    E = np.random.default_rng(1).uniform(1e9, 3e11, len(X))
    nu = np.random.default_rng(2).uniform(0.1, 0.45, len(X))
    Ke = np.stack([element_stiffness_c3d10(X[i], E[i], nu[i]) for i in range(3)])
    from wnfea_amd.elements import c3d10
    Kb = c3d10.element_stiffness_c3d10(X, E, nu)
    for i in range(len(X)):
        Ko = O.element_stiffness(X[i], E[i], nu[i])
        assert np.abs(Kb[i] - Ko).max() <= 1e-12 * np.abs(Ko).max()
        assert np.array_equal(Kb[i], Kb[i].T)
    assert np.allclose(Ke, Kb[:3], rtol=0, atol=0)


def test_ke_matrix_core_variant_matches_oracle():
    """The matrix-core Ke kernel (v_mfma_f64_16x16x4f64 contraction) gives the
    oracle's Ke to FP64 rounding, exactly symmetric, like the vector-ALU
    kernel; the process default is restored afterwards."""
    from oracle import oracle as O
    from wnfea_amd import _lib
    from wnfea_amd.elements import c3d10
    L = _lib.load()
    X = _random_tets(320, seed=7)
    # the midside jitter inverts the odd small element at a Gauss point: keep
    # the valid ones, 301 of them (not a multiple of the 4 elements per block)
    a, b = (5 - 5 ** 0.5) / 20, (5 + 3 * 5 ** 0.5) / 20
    gauss = [(a, a, a), (b, a, a), (a, b, a), (a, a, b)]
    ok = [i for i in range(len(X)) if min(np.linalg.det(O.shape_derivs(*p)[1] @ X[i]) for p in gauss) > 0]
    X = X[ok[:301]]
    assert len(X) == 301
    E = np.random.default_rng(8).uniform(1e9, 3e11, len(X))
    nu = np.random.default_rng(9).uniform(0.1, 0.45, len(X))
    K0 = c3d10.element_stiffness_c3d10(X, E, nu)
    try:
        assert L.wnfea_c3d10_ke_variant(1) == 0
        K1 = c3d10.element_stiffness_c3d10(X, E, nu)
    finally:
        assert L.wnfea_c3d10_ke_variant(0) == 0
    assert L.wnfea_c3d10_ke_variant(2) < 0
    for i in range(len(X)):
        Ko = O.element_stiffness(X[i], E[i], nu[i])
        assert np.abs(K1[i] - Ko).max() <= 1e-12 * np.abs(Ko).max(), i
        assert np.array_equal(K1[i], K1[i].T)
    assert np.abs(K1 - K0).max() <= 1e-12 * np.abs(K0).max()


def test_fint_and_stress_match_oracle():
    from oracle import oracle as O
    from wnfea_amd.elements import c3d10
    X = _random_tets(64, seed=3)
    rng = np.random.default_rng(4)
    U = rng.standard_normal((64, 30)) * 1e-3
    f = c3d10.element_internal_forces_c3d10(X, U, 2.1e11, 0.3)
    s, vm, _ = c3d10.element_stresses_c3d10(X, U, 2.1e11, 0.3)
    for i in range(64):
        fo = O.element_internal_forces(X[i], U[i], 2.1e11, 0.3)
        assert np.linalg.norm(f[i] - fo) <= 1e-12 * np.linalg.norm(fo)
        so, vo, _ = O.element_stresses(X[i], U[i], 2.1e11, 0.3)
        assert np.abs(s[i] - so).max() <= 1e-10 * np.abs(so).max()
        assert np.abs(vm[i] - vo).max() <= 1e-10 * np.abs(vo).max()


def test_inverted_element_raises():
    from wnfea_amd.elements import c3d10
    X = _random_tets(1, seed=5)[0]
    X[:, 0] *= -1.0  # mirror image: every Gauss point has det J < 0
    with pytest.raises(ValueError):
        c3d10.element_stiffness_c3d10(X, 2.1e11, 0.3)


def test_assembly_pattern_and_values_c0(c0_model, c0_oracle):
    """Block pattern bit-exact with the oracle's CSR-derived pattern; values to FP64 round-off."""
    from oracle import oracle as O
    from wnfea_amd.solver.solid_solver import assemble_c3d10_bsr, assemble_c3d10_sparse_3dof
    _, mr, _ = c0_model
    K, F, _ = c0_oracle
    rp, ci, vals = O.csr_to_bsr(K, 3)
    s = assemble_c3d10_bsr(mr)
    assert np.array_equal(s.K.rowptr, rp) and np.array_equal(s.K.colidx, ci)
    v = s.K.get_values().cpu().numpy()
    assert np.abs(v - vals).max() <= 1e-12 * np.abs(vals).max()
    assert np.array_equal(s.F.cpu().numpy(), F)
    Kc, Fc, fixed = assemble_c3d10_sparse_3dof(mr)
    assert fixed == mr.fixed_dof_list(3)
    assert abs(Kc - K).max() <= 1e-12 * abs(K).max()
    assert np.array_equal(Fc, F)


def test_spmv_fp64_fp32_c0(c0_oracle):
    import torch
    from oracle import oracle as O
    from wnfea_amd.solver.bsr_matrix import DeviceBSR, csr_to_bsr
    K, _, _ = c0_oracle
    Kb = csr_to_bsr(K, 3)
    x = np.random.default_rng(42).standard_normal(K.shape[0])
    y_ref = O.csr_spmv(K, x)
    for waves in (1, 2, 4):
        for xcd in (0, 1):
            m = DeviceBSR.from_host(Kb, which=3, sigma=1024)
            m.tune(waves, xcd)
            xd = torch.as_tensor(x, device=m.device)
            y = m.spmv(xd).cpu().numpy()
            assert np.linalg.norm(y - y_ref) <= 1e-14 * np.linalg.norm(y_ref) * 10
            y32 = m.spmv(xd.float(), precision="fp32").cpu().numpy()
            assert np.linalg.norm(y32 - y_ref) <= 1e-5 * np.linalg.norm(y_ref)
            y2 = m.spmv(xd, torch.ones_like(xd), alpha=2.0, beta=-1.0).cpu().numpy()
            assert np.linalg.norm(y2 - (2 * y_ref - 1)) <= 1e-13 * np.linalg.norm(y_ref)
            m.close()


@pytest.mark.parametrize("br,bc", [(6, 6), (3, 6), (6, 3)])
def test_spmv_rect_blocks(br, bc):
    import scipy.sparse as sp
    import torch
    from wnfea_amd.solver.bsr_matrix import DeviceBSR
    rng = np.random.default_rng(br * 10 + bc)
    nr, nc = 700, 450
    A = sp.random(nr, nc, density=0.02, random_state=7, format="csr")
    A.data[:] = 1
    rowptr, colidx = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    vals = rng.standard_normal((A.nnz, br * bc))
    m = DeviceBSR(br, bc, nr, nc, rowptr, colidx, sigma=256)
    m.set_values(torch.as_tensor(vals, device=m.device), which=3)
    x = rng.standard_normal(nc * bc)
    y = m.spmv(torch.as_tensor(x, device=m.device)).cpu().numpy()
    Bs = sp.bsr_matrix((vals.reshape(-1, br, bc), colidx, rowptr), shape=(nr * br, nc * bc))
    y_ref = Bs @ x
    assert np.linalg.norm(y - y_ref) <= 1e-13 * np.linalg.norm(y_ref)
    back = m.get_values().cpu().numpy()
    assert np.array_equal(back, vals)
    m.close()


def test_hbm_stream_read_probe_reads_every_value_once():
    """The streaming-read probe reads the whole buffer exactly once: with
    small integer values every per-thread sum is exact in FP32, so the sums
    over all threads add up to the buffer's sum exactly (one grid below and
    one above the buffer's 16-byte chunk count); bad sizes and alignment are
    refused on the host."""
    import torch
    from wnfea_amd import _lib
    L = _lib.load()
    n = (1 << 20) + 12  # a multiple of four floats, not of the unrolled round
    a = torch.randint(0, 4, (n,), device="cuda", dtype=torch.int32).float()
    ref = float(a.double().sum())
    for grid in (37, 1024):
        out = torch.empty(grid * 512, dtype=torch.float32, device="cuda")
        _lib.check(L.wnfea_hbm_stream_read(_lib.dptr(a), 4 * n, _lib.dptr(out), grid, _lib.stream_of()),
                   "wnfea_hbm_stream_read")
        torch.cuda.synchronize()
        assert float(out.double().sum()) == ref, grid
    out = torch.empty(512, dtype=torch.float32, device="cuda")
    assert L.wnfea_hbm_stream_read(_lib.dptr(a), 4 * n - 4, _lib.dptr(out), 1, _lib.stream_of()) != 0
    assert L.wnfea_hbm_stream_read(_lib.dptr(a) + 4, 4 * n - 16, _lib.dptr(out), 1, _lib.stream_of()) != 0
