"""Adjoint sensitivities of C3D10 responses with respect to per-element
Young's modulus, on the device.

K(E) = sum_e E_e Khat_e (each element stiffness is linear in its own
modulus), so for a response J = c^T u with K u = f:

    dJ/dE_e = -lam_e^T (K_e / E_e) u_e,   K lam = c.

Compliance f^T u is self-adjoint (lam = u: no extra solve); the tip
displacement (mean z-displacement of the tip nodes, the measure of the
reference's I-beam benchmark, scratch/benchmark_c3d10_ibeam.py:139-143)
needs one extra solve.  Fixed DOFs carry u = lam = 0, so the Dirichlet
rows/columns of K do not enter the element products.  The element products
run in one kernel: matrix-free K_e u_e (the internal-force kernel) and the
dot with lam_e, divided by E_e (wnfea_adjoint_grad).  The reference has no
sensitivities; this follows its element and solver conventions.

``C3D10Adjoint`` keeps one assembled system and one SA-AMG hierarchy; new
moduli are assembled in place into the same pattern, and the hierarchy is
rebuilt only when the moduli moved by more than ``rebuild_rel``.  With
``operator="matrix_free"`` (default) the solves run CG on the matrix-free
operator and the V-cycle's finest smoothing steps on it, as the fastest
forward solve does.
``compliance(E)`` and ``tip_displacement(E)`` are torch.autograd functions
of the modulus tensor: their forward runs the forward solve only; the
adjoint solve (tip displacement) and the element sensitivities run in
backward, so a caller under ``torch.no_grad()`` pays for neither.
"""
from __future__ import annotations

import numpy as np

from .. import _lib
from ..model import FEAModel


def element_sensitivities(system, u, lam, E=None):
    """g_e = -lam_e^T (K_e / E_e) u_e for every element (device FP64 tensor)."""
    import torch
    L = _lib.load()
    mesh = system.mesh
    M = mesh["n_elements"]
    E = mesh["E"] if E is None else torch.as_tensor(E, dtype=torch.float64, device=u.device).contiguous()
    u = u.to(torch.float64).contiguous()
    lam = lam.to(torch.float64).contiguous()
    if u.numel() != 3 * mesh["n_nodes"] or lam.numel() != 3 * mesh["n_nodes"]:
        raise ValueError("u and lam must have 3 DOF per node")
    fe = torch.empty((M, 30), dtype=torch.float64, device=u.device)
    grad = torch.empty(M, dtype=torch.float64, device=u.device)
    _lib.check(L.wnfea_adjoint_grad(M, _lib.dptr(mesh["conn"]), _lib.dptr(mesh["nodes"]), _lib.dptr(E),
                                    _lib.dptr(mesh["nu"]), _lib.dptr(u), _lib.dptr(lam), _lib.dptr(fe),
                                    _lib.dptr(grad), _lib.stream_of()), "wnfea_adjoint_grad")
    return grad


def tip_vector(model: FEAModel, tip_nodes, component: int = 2):
    """c with c^T u = mean of u[component] over the tip nodes (zero on fixed DOFs)."""
    n = len(model.mesh_nodes)
    c = np.zeros(3 * n)
    tip = np.asarray(tip_nodes, dtype=np.int64)
    c[3 * tip + component] = 1.0 / max(len(tip), 1)
    c[model.fixed_dof_mask(3)] = 0.0
    return c


class C3D10Adjoint:
    """Forward solves and adjoint gradients of one C3D10 model on the device."""

    def __init__(self, model: FEAModel, tol: float = 1e-10, amg_precision: str = "fp32", rebuild_rel: float = 0.05,
                 amg_kw: dict | None = None, operator: str = "matrix_free"):
        import torch
        from .solid_solver import assemble_c3d10_bsr
        self.model = model
        self.tol = tol
        self.amg_precision = amg_precision
        self.rebuild_rel = rebuild_rel
        self.amg_kw = dict(amg_kw or {})
        self.system = assemble_c3d10_bsr(model, want_fp64_equilibrated=(amg_precision == "fp64"), keep_plan=True)
        self.E_amg = None
        self.amg = None
        self.E_now = self.system.mesh["E"].clone()
        self.solves = 0
        self._torch = torch
        if operator not in ("matrix_free", "assembled"):
            raise ValueError(f"unknown operator {operator!r}")
        self.op = None
        if operator == "matrix_free":
            from .solid_solver import MatrixFreeOperator
            self.op = MatrixFreeOperator(self.system)  # reads the moduli in place

    def _drop_amg(self):
        if self.amg is not None:
            self.amg.close()
            self.amg = None

    def close(self):
        """Free the hierarchy and the matrix-free operator."""
        self._drop_amg()
        if self.op is not None:
            self.op.close()
            self.op = None

    # ------------------------------------------------------------ forward
    def set_moduli(self, E):
        """Assemble K(E) in place (E: M values); rebuild the AMG hierarchy when
        E moved by more than rebuild_rel relative to the one it was built for."""
        from .solid_solver import build_amg, reassemble_c3d10_values
        torch = self._torch
        E = torch.as_tensor(E, dtype=torch.float64, device=self.system.K.device).reshape(-1).contiguous()
        changed = not torch.equal(E, self.E_now)
        if changed:
            reassemble_c3d10_values(self.system, E=E)
            self.E_now = E.clone()
        if self.amg is None or float(((E - self.E_amg).abs() / self.E_amg).max()) > self.rebuild_rel:
            self._drop_amg()
            storage = {"fp64": ("fp64",), "fp32": ("fp32",)}.get(self.amg_precision, ("fp32", "fp16"))
            self.amg = build_amg(self.system, storage=storage, **self.amg_kw)
            self.E_amg = E.clone()
            changed = True
        if changed and self.op is not None:
            self.op.attach_smoother(self.amg)  # the scaling moved with E

    def solve(self, b=None):
        """u = K^-1 b (b = the load vector by default), FP64-verified to tol."""
        from .solid_solver import solve_mixed
        torch = self._torch
        sysm = self.system
        if b is None:
            u, _, _, rel, _ = solve_mixed(sysm, self.amg, tol=self.tol, amg_precision=self.amg_precision,
                                          max_restarts=6, operator=self.op)
        else:
            saved = sysm.F
            try:
                sysm.F = torch.as_tensor(b, dtype=torch.float64, device=sysm.K.device).contiguous()
                u, _, _, rel, _ = solve_mixed(sysm, self.amg, tol=self.tol, amg_precision=self.amg_precision,
                                              max_restarts=6, operator=self.op)
            finally:
                sysm.F = saved
        self.solves += 1
        self.last_rel = rel
        return u

    # ------------------------------------------------------------ responses
    def compliance_value(self, E):
        """(C, u): C = f^T u(E), evaluated as the adjoint-corrected
        2 f^T u - u^T K u, whose error is quadratic in the solve error."""
        self.set_moduli(E)
        u = self.solve()
        F = self.system.F
        return float(2.0 * (F @ u) - u @ self.system.K.spmv(u)), u

    def compliance_and_grad(self, E):
        """C = f^T u and dC/dE (compliance is self-adjoint: lam = u)."""
        C, u = self.compliance_value(E)
        return C, element_sensitivities(self.system, u, u), u

    def response_value(self, E, c):
        """(J, u): J = c^T u(E) for a load-independent output vector c."""
        torch = self._torch
        self.set_moduli(E)
        c = torch.as_tensor(c, dtype=torch.float64, device=self.system.K.device).contiguous()
        u = self.solve()
        return float(c @ u), u

    def response_grad(self, E, c, u):
        """(dJ/dE, lam) for J = c^T u: one adjoint solve K lam = c at the
        moduli E the forward solution u belongs to."""
        torch = self._torch
        self.set_moduli(E)
        c = torch.as_tensor(c, dtype=torch.float64, device=self.system.K.device).contiguous()
        lam = self.solve(c)
        return element_sensitivities(self.system, u, lam), lam

    def response_and_grad(self, E, c):
        """J = c^T u and dJ/dE (one forward and one adjoint solve); J carries
        the adjoint correction lam^T (f - K u)."""
        _, u = self.response_value(E, c)
        g, lam = self.response_grad(E, c, u)
        c = self._torch.as_tensor(c, dtype=self._torch.float64, device=self.system.K.device)
        J = float(c @ u + lam @ (self.system.F - self.system.K.spmv(u)))
        return J, g, u, lam

    def compliance_central_difference(self, E, k: int, h: float, tol: float = 1e-12):
        """(C(E + h e_k) - C(E - h e_k)) / (2 h) from forward solves of the two
        perturbed systems K(E +- h e_k) u+- = f.

        Each perturbed solution is written u+- = u0 + d+- around the solution
        u0 of K(E), so that d+- solves K(E +- h e_k) d = r0 - (K(E +- h e_k) -
        K(E)) u0 with r0 = f - K(E) u0, and C+ - C- = f^T (d+ - d-).  Only the
        rows of element k differ between K(E +- h e_k) and K(E), so the
        right-hand sides are formed without the cancellation of f - K u over
        the whole mesh, and the difference of the two compliances is not the
        small difference of two large numbers.  The perturbed systems are
        assembled on the device (the same reassembly as set_moduli)."""
        torch = self._torch
        E = np.asarray(E.detach().cpu() if isinstance(E, torch.Tensor) else E, dtype=np.float64).copy()
        saved_tol = self.tol
        try:
            self.tol = tol
            self.set_moduli(E)
            u0 = self.solve()
            F = self.system.F
            K0u = self.system.K.spmv(u0)
            r0 = F - K0u
            vals = []
            for sgn in (1.0, -1.0):
                Ep = E.copy()
                Ep[k] += sgn * h
                self.set_moduli(Ep)
                b = r0 - (self.system.K.spmv(u0) - K0u)
                d = self.solve(b)
                vals.append(float(F @ d))
        finally:
            self.tol = saved_tol
            self.set_moduli(E)
        return (vals[0] - vals[1]) / (2.0 * h)

    def compliance(self, E):
        """Compliance f^T u(E) as a differentiable function of the modulus
        tensor: the forward solve runs here, the element sensitivities only
        when the graph is differentiated."""
        return _Compliance.apply(E, self)

    def tip_displacement(self, E, tip_nodes, component: int = 2):
        """Mean tip displacement u(E) as a differentiable function of the
        moduli: the forward solve runs here, the adjoint solve and the
        element sensitivities only when the graph is differentiated."""
        c = tip_vector(self.model, tip_nodes, component)
        return _Response.apply(E, self, c)


def _autograd():
    import torch

    class Compliance(torch.autograd.Function):
        @staticmethod
        def forward(ctx, E, problem):
            C, u = problem.compliance_value(E.detach())
            ctx.problem, ctx.u = problem, u
            ctx.save_for_backward(E)
            return torch.as_tensor(C, dtype=E.dtype, device=E.device)

        @staticmethod
        def backward(ctx, go):
            (E,) = ctx.saved_tensors
            problem = ctx.problem
            problem.set_moduli(E.detach())
            g = element_sensitivities(problem.system, ctx.u, ctx.u)
            return go * g.to(device=go.device, dtype=go.dtype), None

    class Response(torch.autograd.Function):
        @staticmethod
        def forward(ctx, E, problem, c):
            J, u = problem.response_value(E.detach(), c)
            ctx.problem, ctx.u, ctx.c = problem, u, c
            ctx.save_for_backward(E)
            return torch.as_tensor(J, dtype=E.dtype, device=E.device)

        @staticmethod
        def backward(ctx, go):
            (E,) = ctx.saved_tensors
            g, _ = ctx.problem.response_grad(E.detach(), ctx.c, ctx.u)
            return go * g.to(device=go.device, dtype=go.dtype), None, None

    return Compliance, Response


class _Lazy:
    def __init__(self, idx):
        self.idx = idx
        self.fn = None

    def apply(self, *a):
        if self.fn is None:
            self.fn = _autograd()[self.idx]
        return self.fn.apply(*a)


_Compliance = _Lazy(0)
_Response = _Lazy(1)
