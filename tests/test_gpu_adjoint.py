"""Discrete adjoint of the Euler path: vector-Jacobian products of the residual, the RK
right-hand side and the checkpointed SSP-RK3 trajectory, checked by dot-product tests
against central finite differences (h = 1e-6, FP64) of the CPU oracle's residual, right-hand
side, RK3 trajectory and kinetic-energy functional.

Tolerances: the finite differences carry O(h^2) truncation and O(eps/h) rounding, about 1e-8
relative here; the tests accept 1e-6 relative on each directional derivative.
"""
import numpy as np
import pytest

import dgx
import refcpu as R

pytestmark = pytest.mark.gpu
H = 1e-6


def kink_margin(ora, u, basis="dubiner"):
    """Smallest of |u.n| on either side and |s_L - s_R| (s = |u.n| + c) over the interior face
    points: the distance of the state from the kinks of the Rusanov flux."""
    p = ora.order
    s, _ = R.gauss_legendre(p + 1)
    phiF = np.array([[R.basis_eval(basis, p, *R.face_point(f, si))[0] for si in s]
                     for f in range(3)])
    nq = len(s)
    m = np.inf
    for (eL, fL, eR, fR, rev) in ora.interior_faces():
        n, _ = ora.face_geom(eL, fL)
        UL = phiF[fL] @ u[eL]
        UR = (phiF[fR] @ u[eR])[::-1] if rev else phiF[fR] @ u[eR]
        assert UL.shape == (nq, 4)
        sp = []
        for Uq in (UL, UR):
            r = Uq[:, 0]
            un = (Uq[:, 1] * n[0] + Uq[:, 2] * n[1]) / r
            pr = 0.4 * (Uq[:, 3] - 0.5 * (Uq[:, 1] ** 2 + Uq[:, 2] ** 2) / r)
            m = min(m, np.abs(un).min())
            sp.append(np.abs(un) + np.sqrt(1.4 * pr / r))
        m = min(m, np.abs(sp[0] - sp[1]).min())
    return m


def _setup(n, p, precision="fp64", basis="dubiner", seed=1, margin=1e-4):
    verts, tris = R.split_quad_mesh(n)
    ora = R.Oracle(verts, tris, p, basis=basis)
    ctx = dgx.Context(verts, tris, p, basis=basis, precision=precision)
    xy = ora.volume_points()
    # The Rusanov flux has kinks at u.n = 0 and at s_L = s_R; a central difference whose
    # stencil straddles one returns the mean of the one-sided slopes.  The free stream (1, 0)
    # puts every horizontal face on u.n = 0, so the state adds a transverse velocity 0.3 and
    # moves the centre off the mesh lines, and the perturbation seed is advanced until no face
    # point lies within `margin` of a kink (a step of 1e-6 along a unit-variance direction then
    # stays on one side).
    for trial in range(50):
        pert = np.repeat(0.01 * R.uniform_pm1(seed + 1000 * trial, np.arange(ora.K))[:, None],
                         ora.Nq, axis=1)
        U = R.vortex_state(xy, xc=4.71, yc=5.29, rho_pert=pert)
        rho, u, v = U[..., 0], U[..., 1] / U[..., 0], U[..., 2] / U[..., 0] + 0.3
        pr = 0.4 * (U[..., 3] - 0.5 * (U[..., 1] ** 2 + U[..., 2] ** 2) / rho)
        U = np.stack([rho, rho * u, rho * v, pr / 0.4 + 0.5 * rho * (u * u + v * v)], axis=-1)
        u0 = ora.project(U)
        if kink_margin(ora, u0, basis) > margin:
            return ora, ctx, u0
    raise AssertionError("no perturbation seed keeps the state off the flux kinks")


def _dirs(shape, n, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape) for _ in range(n)]


# The Rusanov speed max(s_L, s_R) has a kink wherever s_L = s_R.  With ~1e3 face points and
# face jumps of ~1e-2, a central difference of width h = 1e-6 straddles about one near-tie per
# random direction, and there it returns the mean of the two one-sided slopes instead of the
# derivative.  A direction therefore passes if the central difference at any of
# h = 1e-6, 1e-7, 1e-8 agrees with the adjoint (rounding of the h = 1e-8 difference is ~1e-7
# relative in FP64).
STEPS = (1e-6, 1e-7, 1e-8)


def _agrees(f, u0, d, lam, gd, tol=1e-6):
    errs = []
    for h in STEPS:
        fd = np.sum((f(u0 + h * d) - f(u0 - h * d)) / (2 * h) * lam)
        errs.append(abs(gd - fd) / abs(fd))
        if errs[-1] < (tol if h > 1e-8 else 3 * tol):
            return True, errs
    return False, errs


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_residual_vjp_against_oracle_fd(p):
    ora, ctx, u0 = _setup(5, p)
    rng = np.random.default_rng(p)
    lam = rng.standard_normal(u0.shape)
    g = ctx.residual_vjp(lam, u0)
    for d in _dirs(u0.shape, 8, 10 + p):
        ok, errs = _agrees(ora.residual, u0, d, lam, np.sum(g * d))
        assert ok, errs


def test_rhs_vjp_64_directions():
    # 64 seeded random directions, central differences of the FP64 oracle's M^-1 R
    ora, ctx, u0 = _setup(6, 3)
    lam = np.random.default_rng(2).standard_normal(u0.shape)
    g = ctx.rhs_vjp(lam, u0)
    first = []
    for d in _dirs(u0.shape, 64, 77):
        ok, errs = _agrees(ora.rhs, u0, d, lam, np.sum(g * d))
        assert ok, errs
        first.append(errs[0])
    # at h = 1e-6 alone most directions already agree to 1e-6
    assert np.median(first) < 1e-6, np.median(first)


def test_residual_vjp_monomial_basis():
    ora, ctx, u0 = _setup(4, 2, basis="monomial")
    lam = np.random.default_rng(5).standard_normal(u0.shape)
    g = ctx.residual_vjp(lam, u0)
    d = np.random.default_rng(6).standard_normal(u0.shape)
    ok, errs = _agrees(ora.residual, u0, d, lam, np.sum(g * d))
    assert ok, errs


def test_kinetic_energy_gradient_against_oracle_fd():
    ora, ctx, u0 = _setup(6, 3)
    ctx.set_state(u0)
    J, g = ctx.kinetic_energy_grad()
    assert abs(J - ora.kinetic_energy(u0)) < 1e-12 * abs(J)
    for d in _dirs(u0.shape, 4, 3):
        fd = (ora.kinetic_energy(u0 + H * d) - ora.kinetic_energy(u0 - H * d)) / (2 * H)
        assert abs(np.sum(g * d) - fd) < 1e-7 * abs(fd)


@pytest.mark.parametrize("nsteps,every", [(7, 3), (6, 6), (5, 1)])
def test_rk3_trajectory_vjp_dot_product(nsteps, every):
    # <du_T(u0) d, lam_T> by central differences of the oracle trajectory vs <d, lam_0>
    ora, ctx, u0 = _setup(6, 2)
    ctx.set_state(u0)
    dt = dgx.stable_dt(ctx, 0.3)
    lam_T = np.random.default_rng(4).standard_normal(u0.shape)
    lam0 = ctx.rk3_vjp(dt, nsteps, every, lam_T)
    # the context now holds u_T: compare with the oracle forward
    uT, _ = ora.rk3(u0, dt, nsteps)
    assert np.abs(ctx.get_state() - uT).max() < 1e-11 * np.abs(uT).max()
    traj = lambda u: ora.rk3(u, dt, nsteps)[0]  # noqa: E731
    for d in _dirs(u0.shape, 3, 8):
        ok, errs = _agrees(traj, u0, d, lam_T, np.sum(lam0 * d))
        assert ok, errs


def test_kinetic_energy_adjoint_gradient_against_oracle_fd():
    # gradient of J = int KE at t_final w.r.t. the initial coefficients (the adjoint workload,
    # scaled down): dJ/du0 . d vs central differences of the oracle's J(u_T(u0))
    ora, ctx, u0 = _setup(8, 3, margin=0.0)
    ctx.set_state(u0)
    dt = dgx.stable_dt(ctx, 0.3)
    nsteps = 12
    grad = ctx.rk3_vjp(dt, nsteps, 5)
    J = lambda u: np.array(ora.kinetic_energy(ora.rk3(u, dt, nsteps)[0]))  # noqa: E731
    for d in _dirs(u0.shape, 4, 21):
        ok, errs = _agrees(J, u0, d, np.array(1.0), np.sum(grad * d))
        assert ok, errs


def test_fp32_adjoint_close_to_fp64():
    ora, c64, u0 = _setup(8, 3, "fp64", margin=0.0)
    verts, tris = R.split_quad_mesh(8)
    c32 = dgx.Context(verts, tris, 3, basis="dubiner", precision="fp32")
    c64.set_state(u0)
    c32.set_state(u0)
    dt = dgx.stable_dt(c64, 0.3)
    g64 = c64.rk3_vjp(dt, 10, 4)
    g32 = c32.rk3_vjp(dt, 10, 4)
    assert np.linalg.norm(g32 - g64) / np.linalg.norm(g64) < 1e-4


def test_adjoint_profile_runs_and_restores_state():
    ora, ctx, u0 = _setup(16, 3, "fp32", margin=0.0)
    ctx.set_state(u0)
    before = ctx.get_state()
    dt = dgx.stable_dt(ctx, 0.3)
    f, a, r = ctx.rk3_vjp_profile(dt, 20, 5)
    assert f > 0 and a > 0 and r > 0
    assert np.array_equal(ctx.get_state(), before)


def test_adjoint_rejects_navier_stokes():
    verts, tris = R.split_quad_mesh(3)
    ctx = dgx.Context(verts, tris, 2, weak_form=dgx.NavierStokesWeakFormulation(), precision="fp64")
    with pytest.raises(dgx.DGXError) as e:
        ctx.rhs_vjp(np.zeros((ctx.K, ctx.Np, 4)), np.zeros((ctx.K, ctx.Np, 4)) + 1.0)
    assert e.value.status == dgx.DGX_E_UNSUPPORTED
