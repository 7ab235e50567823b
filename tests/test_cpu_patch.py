"""Patch layout of the matrix-free C3D10 operator, on the host (no device
call): the Morton order is a permutation sorted by the interleaved centroid
code, body by body; every patch's local nodes, 16-bit local connectivity, slot lists and
each node's partial list are consistent with the mesh; and the two kernels'
arithmetic, emulated in numpy with the oracle's element matrices (element
forces summed per local node in slot order, then per node in patch order,
Dirichlet rule applied), reproduces the oracle's assembled K u."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O


def _lib():
    from wnfea_amd import _lib
    return _lib.load()


def _ptr(a):
    return a.ctypes.data


def _box(tets_per_hex, nx=12, ny=3, nz=3):
    from wnfea_amd.mesh.solid_generator import generate_c3d10_structured_block
    m, _ = generate_c3d10_structured_block(1.0, 0.1, 0.1, nx, ny, nz, tip_load_total=-1000.0,
                                           tets_per_hex=tets_per_hex, load_direction="z")
    return m


def _ibeam(layers=4):
    from wnfea_amd.mesh.rcm import apply_rcm_to_model
    from wnfea_amd.mesh.solid_generator import generate_c3d10_ibeam
    m, _ = generate_c3d10_ibeam(n_layers=layers, length=3.0 * layers / 300)
    return apply_rcm_to_model(m)[0]


def morton_order(model):
    L = _lib()
    conn = np.ascontiguousarray(model.solid_elements, dtype=np.int32)
    X = np.ascontiguousarray(np.asarray(model.mesh_nodes, dtype=np.float64)[:, :3])
    order = np.empty(len(conn), dtype=np.int32)
    assert L.wnfea_host_morton_order(len(conn), _ptr(conn), _ptr(X), _ptr(order)) == 0
    return conn, X, order


def patches(model, elems):
    L = _lib()
    conn, X, order = morton_order(model)
    M, N = len(conn), len(model.mesh_nodes)
    npatch = -(-M // elems)
    lmax = ctypes.c_int(0)
    nl = L.wnfea_host_mf_patch_sizes(M, N, _ptr(conn), _ptr(order), elems, ctypes.byref(lmax))
    assert nl > 0
    stride = (lmax.value + 1 + 3) // 4 * 4
    pnode = np.zeros(npatch * stride, dtype=np.int32)
    poff = np.zeros(npatch * stride, dtype=np.uint16)
    pconn = np.zeros(10 * M, dtype=np.uint16)
    pslot = np.zeros(10 * M, dtype=np.uint16)
    nhead = np.zeros(4 * N, dtype=np.int32)
    novf = np.zeros(nl + 1, dtype=np.int32)
    nov = L.wnfea_host_mf_patches(M, N, _ptr(conn), _ptr(order), elems, stride, _ptr(pnode), _ptr(pconn),
                                  _ptr(poff), _ptr(pslot), _ptr(nhead), _ptr(novf))
    assert nov >= 0
    return dict(conn=conn, X=X, order=order, M=M, N=N, P=elems, npatch=npatch, stride=stride, pnode=pnode,
                pconn=pconn, poff=poff, pslot=pslot, nhead=nhead.reshape(N, 4), novf=novf[:nov], lmax=lmax.value,
                n_local=nl)


def node_partials(t, n):
    """Partial slots of node n in the order the node kernel adds them."""
    h = t["nhead"][n]
    out = [int(v) for v in h[:3] if v >= 0]
    if h[3] < 0:
        k = -int(h[3]) - 1
        out += [int(v) for v in t["novf"][k + 1:k + 1 + t["novf"][k]]]
    else:
        assert h[3] == len(out)
    return out


def test_morton_order_is_the_sorted_interleaved_centroid_code():
    m = _ibeam(4)
    conn, X, order = morton_order(m)
    M = len(conn)
    assert np.array_equal(np.sort(order), np.arange(M))
    c = X[conn[:, :4]].mean(axis=1)
    lo, ext = c.min(0), (c.max(0) - c.min(0)).max()
    q = np.clip(np.floor((c - lo) * ((2 ** 21 - 1) / ext)), 0, 2 ** 21 - 1).astype(np.uint64)
    code = np.zeros(M, dtype=np.uint64)
    for b in range(21):
        for a in range(3):
            code |= ((q[:, a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + a)
    ref = np.lexsort((np.arange(M), code))
    assert np.array_equal(order, ref)


def test_morton_order_keeps_separate_bodies_apart():
    """Two copies of one box in the same place (a batch of meshes in one
    system): the order takes every element of the first copy, then every
    element of the second, each copy in its own Morton order, so no patch
    mixes the two bodies."""
    L = _lib()
    m = _box(6)
    conn = np.ascontiguousarray(m.solid_elements, dtype=np.int32)
    X = np.asarray(m.mesh_nodes, dtype=np.float64)[:, :3]
    M, N = len(conn), len(X)
    one = np.empty(M, dtype=np.int32)
    assert L.wnfea_host_morton_order(M, _ptr(conn), _ptr(np.ascontiguousarray(X)), _ptr(one)) == 0
    conn2 = np.ascontiguousarray(np.vstack([conn, conn + N]), dtype=np.int32)
    X2 = np.ascontiguousarray(np.vstack([X, X]))
    two = np.empty(2 * M, dtype=np.int32)
    assert L.wnfea_host_morton_order(2 * M, _ptr(conn2), _ptr(X2), _ptr(two)) == 0
    assert np.array_equal(two[:M], one) and np.array_equal(two[M:], one + M)


@pytest.mark.parametrize("mesh,elems", [("box6", 64), ("box5", 128), ("ibeam", 128), ("ibeam", 256)])
def test_patch_layout_is_consistent_with_the_mesh(mesh, elems):
    m = {"box6": lambda: _box(6), "box5": lambda: _box(5), "ibeam": lambda: _ibeam(4)}[mesh]()
    t = patches(m, elems)
    conn, order, P, M, N, S = t["conn"], t["order"], t["P"], t["M"], t["N"], t["stride"]
    pnode, pconn, poff, pslot = t["pnode"], t["pconn"], t["poff"], t["pslot"]
    assert S > t["lmax"] and S % 4 == 0
    total = 0
    for p in range(t["npatch"]):
        s0, s1 = p * P, min(M, (p + 1) * P)
        ne = s1 - s0
        row = pnode[p * S:(p + 1) * S]
        nl = int((row >= 0).sum())
        assert np.all(row[:nl] >= 0) and np.all(row[nl:] == -1)
        loc = row[:nl]
        total += nl
        # local nodes: the patch's distinct nodes, ascending
        assert np.array_equal(loc, np.unique(conn[order[s0:s1]]))
        # the local connectivity maps back to the mesh
        pc = pconn[10 * s0:10 * s1].astype(np.int64)
        assert np.array_equal(loc[pc], conn[order[s0:s1]].ravel())
        # slot lists: a partition of the patch's slot codes, each list ascending and on its node
        offs = poff[p * S:(p + 1) * S].astype(np.int64)
        assert np.all(offs[nl:] == 10 * ne)
        codes = pslot[10 * s0:10 * s0 + 10 * ne].astype(np.int64)
        assert np.array_equal(np.sort(codes), np.arange(10 * ne))
        for j in range(nl):
            lst = codes[offs[j]:offs[j + 1]]
            assert np.all(np.diff(lst) > 0) and np.all(pc[lst] == j)
    assert total == t["n_local"] and t["lmax"] == max(int((pnode[p * S:(p + 1) * S] >= 0).sum())
                                                       for p in range(t["npatch"]))
    # each node's partials: its local copies, in patch order
    seen = []
    for n in range(N):
        lst = node_partials(t, n)
        seen += lst
        assert np.all(pnode[lst] == n) and np.all(np.diff(lst) > 0)
        assert len(lst) == len(np.unique(np.asarray(lst) // S))
    assert np.array_equal(np.sort(seen), np.flatnonzero(pnode >= 0))


@pytest.mark.parametrize("mesh,elems", [("box5", 64), ("ibeam", 128)])
def test_patch_arithmetic_reproduces_the_assembled_operator(mesh, elems):
    """The kernels' sums in numpy: f_e = K_e (mask .* u)_e, per local node the
    element forces in slot order, per node its partials in patch order, and a
    fixed DOF keeps u (unit diagonal): equals the oracle's assembled K u."""
    m = {"box5": lambda: _box(5), "ibeam": lambda: _ibeam(4)}[mesh]()
    t = patches(m, elems)
    conn, order, P, M, N = t["conn"], t["order"], t["P"], t["M"], t["N"]
    E, nu = m.element_moduli()
    fixed = m.fixed_dof_mask(3)
    K, _ = O.assemble_csr(t["X"], conn, E, nu, fixed, m.load_vector(3))
    Ke = O.ke_batch(t["X"], conn, E, nu)
    u = np.random.default_rng(1).standard_normal(3 * N)
    us = np.where(fixed, 0.0, u)
    dofs = (3 * conn[:, :, None] + np.arange(3)).reshape(M, 30)
    fe = np.einsum("eij,ej->ei", Ke, us[dofs]).reshape(M, 10, 3)
    S = t["stride"]
    part = np.zeros((t["npatch"] * S, 3))
    pslot, poff = t["pslot"], t["poff"]
    for p in range(t["npatch"]):
        s0, ne = p * P, min(M, (p + 1) * P) - p * P
        codes = pslot[10 * s0:10 * s0 + 10 * ne].astype(np.int64)
        offs = poff[p * S:(p + 1) * S].astype(np.int64)
        vals = fe[order[s0 + codes // 10], codes % 10]
        for j in range(int((t["pnode"][p * S:(p + 1) * S] >= 0).sum())):
            acc = np.zeros(3)
            for v in vals[offs[j]:offs[j + 1]]:
                acc = acc + v
            part[p * S + j] = acc
    y = np.zeros((N, 3))
    for n in range(N):
        acc = np.zeros(3)
        for k in node_partials(t, n):
            acc = acc + part[k]
        y[n] = acc
    y = y.ravel()
    y[fixed] = u[fixed]
    ref = O.csr_spmv(K, u)
    assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)


def test_patch_statistics_of_the_ibeam_section():
    """The Morton patches of the I-beam section are compact: with 128
    elements per patch over half of the nodes are interior to one patch,
    over 95 % have at most three partial sums (the node head), and the
    largest patch keeps an FP64 workgroup's LDS under 64 KiB."""
    t = patches(_ibeam(8), 128)
    counts = np.array([len(node_partials(t, n)) for n in range(t["N"])])
    assert (counts == 1).mean() > 0.5 and (counts <= 3).mean() > 0.95
    assert (3 * t["stride"] + 30 * 128) * 8 + 2 * (10 * 128 + t["stride"]) <= 64 * 1024
    assert t["n_local"] / t["M"] < 3.0


def test_patch_layout_rejects_bad_input():
    L = _lib()
    m = _box(6, 4, 2, 2)
    conn, X, order = morton_order(m)
    M, N = len(conn), len(m.mesh_nodes)
    lmax = ctypes.c_int(0)

    def sizes(o, elems=64, c=conn):
        return L.wnfea_host_mf_patch_sizes(M, N, _ptr(c), _ptr(o), elems, ctypes.byref(lmax))
    assert sizes(order) > 0
    stride = lmax.value + 1
    bufs = [np.zeros(-(-M // 64) * stride + 8, dtype=np.int32), np.zeros(10 * M, dtype=np.uint16),
            np.zeros(-(-M // 64) * stride + 8, dtype=np.uint16), np.zeros(10 * M, dtype=np.uint16),
            np.zeros(4 * N, dtype=np.int32), np.zeros(10 * M, dtype=np.int32)]

    def fill(o, s=stride):
        return L.wnfea_host_mf_patches(M, N, _ptr(conn), _ptr(o), 64, s, *[_ptr(b) for b in bufs])
    assert fill(order) >= 0
    assert fill(order, s=lmax.value) == -1      # stride must exceed the most local nodes
    dup = order.copy()
    dup[1] = dup[0]
    assert sizes(dup) == -1 and fill(dup) == -1  # not a permutation
    assert sizes(order, elems=0) == -1
    bad = conn.copy()
    bad[0, 0] = N
    assert sizes(order, c=bad) == -1             # corner outside the mesh
