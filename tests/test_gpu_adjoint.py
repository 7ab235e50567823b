"""Device adjoint sensitivities (compliance and tip displacement with respect
to per-element Young's modulus) against the CPU FP64 reference gradient and
central finite differences, on C0 with seeded random moduli."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3(c0_model):
    from scipy.sparse.linalg import spsolve
    from oracle import oracle as O
    from wnfea_amd.solver.adjoint import C3D10Adjoint, tip_vector
    _, mr, _ = c0_model
    M = len(mr.solid_elements)
    E = 210e9 * (0.2 + 0.8 * np.random.default_rng(2).uniform(size=M))
    nu = np.full(M, 0.3)
    mr = __import__("copy").copy(mr)
    mr.solid_E, mr.solid_nu = E, nu
    tip = np.flatnonzero(np.abs(mr.mesh_nodes[:, 0] - 1.0) < 1e-12)
    c = tip_vector(mr, tip)
    K, F = O.assemble_csr(mr.mesh_nodes, mr.solid_elements, E, nu, mr.fixed_dof_mask(3), mr.load_vector(3))
    u = spsolve(K.tocsc(), F)
    lam = spsolve(K.tocsc(), c)
    ref = {"gC": O.adjoint_grad(mr.mesh_nodes, mr.solid_elements, E, nu, u, u),
           "gT": O.adjoint_grad(mr.mesh_nodes, mr.solid_elements, E, nu, u, lam),
           "C": F @ u, "T": c @ u}
    prob = C3D10Adjoint(mr, tol=1e-11, amg_precision="fp32")
    yield mr, E, tip, c, ref, prob
    prob.close()


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def test_compliance_gradient_vs_cpu_reference(c3):
    import torch
    mr, E, tip, c, ref, prob = c3
    Et = torch.tensor(E, dtype=torch.float64, device="cuda", requires_grad=True)
    C = prob.compliance(Et)
    C.backward()
    assert abs(float(C) - ref["C"]) <= 1e-8 * abs(ref["C"])
    assert _rel(Et.grad.cpu().numpy(), ref["gC"]) <= 1e-5


def test_tip_displacement_gradient_vs_cpu_reference(c3):
    import torch
    mr, E, tip, c, ref, prob = c3
    Et = torch.tensor(E, dtype=torch.float64, device="cuda", requires_grad=True)
    T = prob.tip_displacement(Et, tip)
    (2.0 * T).backward()  # chain rule through the autograd function
    assert abs(float(T) - ref["T"]) <= 1e-8 * abs(ref["T"])
    assert _rel(Et.grad.cpu().numpy(), 2.0 * ref["gT"]) <= 1e-5


def test_gradient_vs_device_central_differences(c3):
    """h = 1e-4 E_e on 5 seeded elements, device forward solves of the
    perturbed systems (same hierarchy, values reassembled in place), with the
    adjoint-corrected compliance."""
    mr, E, tip, c, ref, prob = c3
    rng = np.random.default_rng(3)
    _, g, _ = prob.compliance_and_grad(E)
    g = g.cpu().numpy()
    for e in rng.choice(len(E), 5, replace=False):
        h = 1e-4 * E[e]
        vals = []
        for sgn in (1.0, -1.0):
            Ep = E.copy()
            Ep[e] += sgn * h
            Cp, _, _ = prob.compliance_and_grad(Ep)
            vals.append(Cp)
        fd = (vals[0] - vals[1]) / (2 * h)
        # the FD difference of C carries FP64 round-off of C itself, an absolute
        # error: measure it against the gradient's scale as well as g_e
        assert abs(fd - g[e]) <= 2e-3 * abs(g[e]) + 1e-4 * np.abs(g).max(), (e, fd, g[e])
    prob.set_moduli(E)


def test_matrix_free_and_assembled_operators_agree_after_new_moduli(c3):
    """The adjoint solves run on the matrix-free operator by default; they
    agree with the assembled-operator run, and after new moduli (reassembled
    in place, hierarchy rebuilt) the matrix-free solution satisfies the
    oracle's K(E2), so the operator reads the updated moduli."""
    from oracle import oracle as O
    from wnfea_amd.solver.adjoint import C3D10Adjoint
    mr, E, tip, c, ref, prob = c3
    pa = C3D10Adjoint(mr, tol=1e-11, amg_precision="fp32", operator="assembled")
    try:
        C1, g1, _ = prob.compliance_and_grad(E)
        C2, g2, _ = pa.compliance_and_grad(E)
        assert abs(C1 - C2) <= 1e-9 * abs(C2) and _rel(g1.cpu().numpy(), g2.cpu().numpy()) <= 1e-7
        E2 = E * (1.0 + 0.2 * np.random.default_rng(8).uniform(size=len(E)))
        C1b, g1b, u1 = prob.compliance_and_grad(E2)
        C2b, g2b, _ = pa.compliance_and_grad(E2)
        assert abs(C1b - C2b) <= 1e-9 * abs(C2b) and _rel(g1b.cpu().numpy(), g2b.cpu().numpy()) <= 1e-7
        K2, F = O.assemble_csr(mr.mesh_nodes, mr.solid_elements, E2, np.full(len(E2), 0.3), mr.fixed_dof_mask(3),
                               mr.load_vector(3))
        u1 = u1.cpu().numpy()
        assert np.linalg.norm(F - K2 @ u1) <= 1e-9 * np.linalg.norm(F)
        assert prob.compliance_and_grad(E)[0] == pytest.approx(C1, rel=1e-9)  # back to the fixture's moduli
    finally:
        pa.close()
