"""CPU mirror of the library's V-cycle (test helper, FP64 numpy).

Same steps as csrc/solver.hip on the host hierarchy from sa_setup:
Chebyshev(deg) pre-smoothing from zero, defect, R, recursion, P, Chebyshev
post-smoothing, dense inverse on the coarsest level."""
import numpy as np


def cheb_coefs(lmax, ratio, deg):
    lmin = lmax / ratio
    th, de = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    sig = th / de
    rho = 1.0 / sig
    c = [(0.0, 1.0 / th)]
    for _ in range(1, deg):
        rn = 1.0 / (2.0 * sig - rho)
        c.append((rn * rho, 2.0 * rn / de))
        rho = rn
    return c


def vcycle(levels, inv, b, l=0, deg=2, ratio=30.0):
    """deg: Chebyshev degree, one for all levels or a list per level.  A level
    may carry ``A_smooth``, the operator of its smoothing steps (the defect
    always uses ``A``)."""
    lv = levels[l]
    if l == len(levels) - 1:
        return inv @ b
    A, dinv = lv.A, lv.dinv
    As = getattr(lv, "A_smooth", A)
    dl = deg[l] if isinstance(deg, (list, tuple)) else deg
    c = cheb_coefs(lv.lmax, ratio, dl)
    d = c[0][1] * dinv * b
    x = d.copy()
    for k in range(1, dl):
        d = c[k][0] * d + c[k][1] * dinv * (b - As @ x)
        x = x + d
    res = b - A @ x
    e = vcycle(levels, inv, lv.R @ res, l + 1, deg, ratio)
    x = x + lv.P @ e
    for k in range(dl):
        d = c[k][0] * d + c[k][1] * dinv * (b - As @ x)
        x = x + d
    return x


def pcg(A, b, M, rtol, maxit=1000):
    x = np.zeros_like(b)
    r = b.copy()
    z = M(r)
    p = z.copy()
    rz = r @ z
    nb = np.linalg.norm(b)
    for it in range(maxit):
        q = A @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        if np.linalg.norm(r) <= rtol * nb:
            return x, it + 1
        z = M(r)
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    return x, maxit
