"""Analytic flows of the reference's end-to-end Navier-Stokes accuracy tests and their error
norm, restated for the tests (test data, not product code):

* the manufactured shear W = (1, y, 0, 1) (T/solver/shear_flow_accuracy_test.cpp:18-30);
* the compressible Blasius profile (T/solver/compressible_blasius_accuracy_test.cpp:17-152):
  the Blasius ODE f''' = -f f''/2 integrated by RK4 from f''(0) = 0.332057336215 with step
  0.01 to eta = 12 (f' clamped to 1 once above 0.999999), linear interpolation in eta, the
  Crocco-Busemann temperature with recovery factor Pr^(1/3), constant pressure;
* compute_variable_error (:32-79 of the shear test): the relative L2 error of one primitive
  field with weights w_q |J_q|, absolute when the exact field's norm is below 1e-14.
"""
import bisect
import math

import numpy as np

import refcpu as R

GAMMA = 1.4


def prim2cons(W, g=GAMMA):
    rho, u, v, p = W
    return [rho, rho * u, rho * v, p / (g - 1) + 0.5 * rho * (u * u + v * v)]


def shear_primitive(x, y):
    return [1.0, y, 0.0, 1.0]


def _blasius_table():
    d = 0.01
    n = int(12.0 / d) + 1
    eta, f, fp, fpp = 0.0, 0.0, 0.0, 0.332057336215
    E, F, FP = [], [], []

    def rhs(a, b, c):
        return b, c, -0.5 * a * c

    for _ in range(n):
        E.append(eta)
        F.append(f)
        FP.append(fp)
        k1 = rhs(f, fp, fpp)
        k2 = rhs(f + 0.5 * d * k1[0], fp + 0.5 * d * k1[1], fpp + 0.5 * d * k1[2])
        k3 = rhs(f + 0.5 * d * k2[0], fp + 0.5 * d * k2[1], fpp + 0.5 * d * k2[2])
        k4 = rhs(f + d * k3[0], fp + d * k3[1], fpp + d * k3[2])
        f += d / 6.0 * (k1[0] + 2.0 * k2[0] + 2.0 * k3[0] + k4[0])
        fp += d / 6.0 * (k1[1] + 2.0 * k2[1] + 2.0 * k3[1] + k4[1])
        fpp += d / 6.0 * (k1[2] + 2.0 * k2[2] + 2.0 * k3[2] + k4[2])
        eta += d
        if fp > 0.999999:
            fp, fpp = 1.0, 0.0
    return E, F, FP


_TABLE = None


def _interp(xs, ys, x):
    if x <= xs[0]:
        return ys[0]
    if x >= xs[-1]:
        return ys[-1]
    i = bisect.bisect_right(xs, x)
    t = (x - xs[i - 1]) / (xs[i] - xs[i - 1])
    return ys[i - 1] + t * (ys[i] - ys[i - 1])


# the Blasius test's parameters (:205-214)
BLASIUS = dict(mu=5.0e-5, prandtl=0.72, mach=0.3, T_inf=1.0, T_wall=1.15, p_inf=1.0, x_min=0.05)


def blasius_primitive(x, y, g=GAMMA):
    global _TABLE
    if _TABLE is None:
        _TABLE = _blasius_table()
    E, F, FP = _TABLE
    b = BLASIUS
    xp = max(x, b["x_min"])
    nu = b["mu"] / (b["p_inf"] / b["T_inf"])
    U = b["mach"] * math.sqrt(g * b["T_inf"])
    eta = min(max(y * math.sqrt(U / (2.0 * nu * xp)), E[0]), E[-1])
    f, fp = _interp(E, F, eta), _interp(E, FP, eta)
    T = b["T_inf"] * (1.0 + b["prandtl"] ** (1.0 / 3.0) * (b["T_wall"] / b["T_inf"] - 1.0) * fp
                      + 0.5 * (g - 1.0) * b["mach"] ** 2 * (1.0 - fp * fp))
    return [b["p_inf"] / T, U * fp, 0.5 * math.sqrt(nu * U / xp) * (eta * fp - f), b["p_inf"]]


def volume_weights(ora):
    if ora.nvpe == 4:
        _, w = R.gauss_legendre(ora.order + 1)
        return np.outer(w, w).ravel()
    _, _, w = R.quad_tri(2 * ora.order)
    return w


def l2_errors(ora, u, prim_fn, g=GAMMA):
    """compute_variable_error for the four primitive fields of the state u."""
    xy = ora.volume_points()
    return l2_errors_at_points(ora, u, np.array([[prim_fn(*q) for q in e] for e in xy]), g)


def l2_errors_at_points(ora, u, Wex, g=GAMMA):
    """The same with the exact primitive values given at the oracle's volume points."""
    Uq = ora.eval_volume(u)
    w = volume_weights(ora)
    dw = np.array([w * np.abs(ora.elem_data(e)[0]) for e in range(ora.K)])[..., None]
    rho = Uq[..., 0]
    W = np.stack([rho, Uq[..., 1] / rho, Uq[..., 2] / rho,
                  (g - 1) * (Uq[..., 3] - 0.5 * (Uq[..., 1] ** 2 + Uq[..., 2] ** 2) / rho)], axis=-1)
    e2, n2 = ((W - Wex) ** 2 * dw).sum(axis=(0, 1)), (Wex ** 2 * dw).sum(axis=(0, 1))
    return np.where(n2 < 1e-14, np.sqrt(e2), np.sqrt(e2 / np.where(n2 < 1e-14, 1.0, n2)))


def ns_oracle(verts, elems, order, mu, sigma0, prim_fn, points="reference", **kw):
    """NS oracle with the far-field condition prim2cons(prim_fn(x)) on every tag, and the
    projected initial state."""
    ora = R.Oracle(verts, elems, order, periodic=0, physics="ns", mu=mu, prandtl=0.72, sigma0=sigma0,
                   ns_face_grad="reference", euler_bc_points=points, **kw)
    for tag in ("Bottom", "Right", "Top", "Left"):
        ora.set_bc(tag, "far_field", fn=lambda x, y: prim2cons(prim_fn(x, y)))
    xy = ora.volume_points()
    u0 = ora.project(np.array([[prim2cons(prim_fn(*q)) for q in e] for e in xy]))
    return ora, u0
