"""wnfea_residual_noalloc -- the residual entry point that allocates nothing
(reduction partials in the library's static per-GPU scratch) -- against
wnfea_residual (partials in a solver workspace), through the C-ABI: the same
r = f - K u bit for bit and the same two norms, on two calls in a row, the
second on a smaller system than the first."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def _box(nx, ny, nz):
    from wnfea_amd.mesh.rcm import apply_rcm_to_model
    from wnfea_amd.mesh.solid_generator import generate_c3d10_structured_block
    from wnfea_amd.solver.solid_solver import assemble_c3d10_bsr
    m, _ = generate_c3d10_structured_block(1.0, 0.1, 0.1, nx, ny, nz, tip_load_total=-1000.0, tets_per_hex=6,
                                           load_direction="z")
    m, _, _ = apply_rcm_to_model(m)
    return assemble_c3d10_bsr(m)


def test_noalloc_residual_equals_workspace_residual_twice():
    import torch
    from wnfea_amd import _lib
    from wnfea_amd.solver.solid_solver import _Solver
    L = _lib.load()
    noalloc = L.wnfea_residual_noalloc
    g = torch.Generator(device="cuda").manual_seed(3)
    for dims in ((40, 4, 4), (12, 2, 2)):  # the second call on a smaller system
        system = _box(*dims)
        n = system.n_dofs
        u = torch.randn(n, dtype=torch.float64, device="cuda", generator=g) * 1e-6
        r_ref = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        r_new = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        out_ref = (ctypes.c_double * 2)()
        _lib.check(L.wnfea_residual(_Solver.get(), system.K.handle, _lib.dptr(u), _lib.dptr(system.F),
                                    _lib.dptr(r_ref), out_ref, _lib.stream_of()), "wnfea_residual")
        out_dev = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
        _lib.check(noalloc(system.K.handle, _lib.dptr(u), _lib.dptr(system.F), _lib.dptr(r_new), _lib.dptr(out_dev),
                           _lib.stream_of()), "wnfea_residual_noalloc")
        torch.cuda.synchronize()
        assert bool(torch.equal(r_new, r_ref)), dims
        got = out_dev.cpu().numpy()
        assert got[0] == out_ref[0] and got[1] == out_ref[1], (dims, got, list(out_ref))
        # and against a plain FP64 reference of the same quantities
        exp = system.F - system.K.spmv(u)
        assert float(torch.linalg.norm(r_new - exp)) <= 1e-12 * float(torch.linalg.norm(exp))
        assert abs(got[1] - float(torch.linalg.norm(system.F))) <= 1e-12 * got[1]
        system.K.close()


def test_noalloc_residual_without_r_and_bad_args():
    import torch
    from wnfea_amd import _lib
    L = _lib.load()
    noalloc = L.wnfea_residual_noalloc
    system = _box(10, 2, 2)
    u = torch.zeros(system.n_dofs, dtype=torch.float64, device="cuda")
    out_dev = torch.zeros(2, dtype=torch.float64, device="cuda")
    _lib.check(noalloc(system.K.handle, _lib.dptr(u), _lib.dptr(system.F), None, _lib.dptr(out_dev),
                       _lib.stream_of()), "wnfea_residual_noalloc")
    got = out_dev.cpu().numpy()
    fn = float(torch.linalg.norm(system.F))
    assert abs(got[0] - fn) <= 1e-12 * fn and abs(got[1] - fn) <= 1e-12 * fn  # u = 0: r = f
    assert noalloc(None, _lib.dptr(u), _lib.dptr(system.F), None, _lib.dptr(out_dev), _lib.stream_of()) < 0
    assert noalloc(system.K.handle, _lib.dptr(u), _lib.dptr(system.F), None, None, _lib.stream_of()) < 0
    system.K.close()


def test_noalloc_residual_rejects_matrices_beyond_its_scratch():
    """One block row per slice (64 lanes per row): 70 000 slices exceed the
    65 528 the static scratch holds, and the call returns the bad-size
    status without launching anything."""
    import numpy as np
    import torch
    from wnfea_amd import _lib
    from wnfea_amd.solver.bsr_matrix import DeviceBSR
    L = _lib.load()
    n = 70000
    big = DeviceBSR(3, 3, n, n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), sigma=64, lanes=64)
    try:
        assert big.info()["n_slices"] > 65528
        x = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
        out_dev = torch.zeros(2, dtype=torch.float64, device="cuda")
        assert L.wnfea_residual_noalloc(big.handle, _lib.dptr(x), _lib.dptr(x), None, _lib.dptr(out_dev),
                                        _lib.stream_of()) == -2
    finally:
        big.close()
