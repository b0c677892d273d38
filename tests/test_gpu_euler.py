"""GPU parity of the Euler residual path against the FP64 CPU oracle.

Tolerances: FP64 kernels match the oracle to reassociation noise (relative 1e-12 on the
residual, 1e-10 after 100 RK3 steps); FP32 kernels to relative 2e-5 on the residual and
to per-field relative L2 1e-5 after 20 steps.
"""
import numpy as np
import pytest

import dgx
import refcpu as R

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _vortex_setup(n, p, basis, precision, seed=None):
    verts, tris = R.split_quad_mesh(n)
    ora = R.Oracle(verts, tris, p, basis=basis)
    ctx = dgx.Context(verts, tris, p, basis=basis, precision=precision)
    xy = ora.volume_points()
    pert = None
    if seed is not None:
        pert = np.repeat(0.01 * R.uniform_pm1(seed, np.arange(ora.K))[:, None], ora.Nq, axis=1)
    u0 = ora.project(R.vortex_state(xy, rho_pert=pert))
    return ora, ctx, u0


@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("basis", ["monomial", "dubiner"])
def test_residual_fp64_matches_oracle(p, basis):
    ora, ctx, u0 = _vortex_setup(6, p, basis, "fp64", seed=7)
    r_ref = ora.residual(u0)
    r_gpu = ctx.residual(u0)
    assert _rel(r_gpu, r_ref) < 1e-12


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_rhs_fp64_matches_oracle(p):
    ora, ctx, u0 = _vortex_setup(5, p, "dubiner", "fp64", seed=3)
    k_ref = ora.rhs(u0)
    k_gpu = ctx.apply_minv(ctx.residual(u0))
    assert _rel(k_gpu, k_ref) < 1e-12


@pytest.mark.parametrize("p", [2, 3])
def test_rk3_steps_fp64_c1_shape(p):
    # the CPU-runnable configuration: 32x32 split quads (K = 2048), FP64, monomial basis,
    # 100 SSP-RK3 steps at CFL 0.3 (p = 2 is the configuration itself)
    ora, ctx, u0 = _vortex_setup(32, p, "monomial", "fp64")
    ctx.set_state(u0)
    dt = dgx.stable_dt(ctx, 0.3)
    u_ref, _ = ora.rk3(u0, dt, 100)
    ctx.rk3_steps(dt, 100)
    u_gpu = ctx.get_state()
    assert _rel(u_gpu, u_ref) < 1e-10


def test_c1_monomial_and_dubiner_agree_on_the_gpu():
    # both bases span the same polynomials: after 100 steps from the same initial condition the
    # point values at the volume quadrature points agree to FP64 roundoff
    vals = []
    dt = None
    for basis in ("monomial", "dubiner"):
        ora, ctx, u0 = _vortex_setup(32, 2, basis, "fp64")
        ctx.set_state(u0)
        dt = dt or dgx.stable_dt(ctx, 0.3)
        ctx.rk3_steps(dt, 100)
        vals.append(ctx.eval_volume())
    assert _rel(vals[1], vals[0]) < 1e-12


@pytest.mark.parametrize("p", [1, 3, 4])
def test_residual_fp32_close_to_oracle(p):
    ora, ctx, u0 = _vortex_setup(6, p, "dubiner", "fp32", seed=11)
    r_ref = ora.residual(u0)
    r_gpu = ctx.residual(u0)
    # FP32 evaluation of O(1) fluxes: relative 2e-5 of the residual's max entry
    assert _rel(r_gpu, r_ref) < 2e-5


def point_permutation(xy_a, xy_b):
    """Index map perm with xy_a[:, k] == xy_b[:, perm[k]]: the product and the oracle expand the
    same quadrature orbits in different orders."""
    d = np.abs(xy_a[0][:, None, :] - xy_b[0][None, :, :]).sum(-1)
    perm = d.argmin(axis=1)
    assert len(set(perm.tolist())) == len(perm)
    assert np.allclose(xy_a, xy_b[:, perm], rtol=1e-14, atol=1e-13)
    return perm


def test_rk3_fp32_field_error():
    ora, ctx, u0 = _vortex_setup(16, 3, "dubiner", "fp32", seed=5)
    ctx.set_state(u0)
    dt = dgx.stable_dt(ctx, 0.3)
    u_ref, _ = ora.rk3(u0, dt, 20)
    ctx.rk3_steps(dt, 20)
    u_gpu = ctx.get_state()
    # Dubiner coefficients are orthonormal (Gram 2I) and the elements have equal areas, so the
    # coefficient 2-norm per variable is the field L2 norm up to a constant
    for v in range(4):
        err = np.linalg.norm(u_gpu[..., v] - u_ref[..., v]) / np.linalg.norm(u_ref[..., v])
        assert err < 1e-5, (v, err)


def test_fp32_state_round_trip_and_diagnostics():
    # the FP32 state is stored as a deviation from its mean on the constant mode: set/get
    # round-trips to FP32 rounding of that deviation, and every diagnostic adds the offset back
    ora, ctx, u0 = _vortex_setup(12, 3, "dubiner", "fp32", seed=9)
    ctx.set_state(u0)
    back = ctx.get_state()
    dev = u0.copy()
    dev[:, 0, :] -= u0[:, 0, :].mean(axis=0)
    # the offset itself is the mean rounded to FP32
    slack = 2 ** -22 * np.abs(u0[:, 0, :].mean(axis=0))
    assert np.all(np.abs(back - u0) <= 2 ** -23 * np.abs(dev) + slack)
    s_ref = ora.conservation_sums(u0)
    assert np.allclose(ctx.conservation_sums(), s_ref, rtol=1e-6, atol=1e-6)
    assert abs(ctx.kinetic_energy() - ora.kinetic_energy(u0)) < 1e-6 * ora.kinetic_energy(u0)
    assert np.allclose(ctx.density_range(), ora.density_range(u0), rtol=1e-6)
    assert abs(ctx.max_cfl(0.01) - ora.max_cfl(u0, 0.01)) < 1e-6 * ora.max_cfl(u0, 0.01)
    xy = ctx.volume_points()
    perm = point_permutation(xy, ora.volume_points())
    assert _rel(ctx.eval_volume(), ora.eval_volume(u0)[:, perm]) < 1e-6
    # a projection sets its own offset (the mean of the point values)
    ctx.project(R.vortex_state(xy))
    assert _rel(ctx.get_state(), ora.project(R.vortex_state(ora.volume_points()))) < 1e-6


def test_rk3_fp32_error_scales_with_deviation():
    # 100 steps at the headline CFL: with the offset the density error stays below 1e-6
    ora, ctx, u0 = _vortex_setup(16, 3, "dubiner", "fp32", seed=5)
    c64 = dgx.Context(*R.split_quad_mesh(16), 3, basis="dubiner", precision="fp64")
    ctx.set_state(u0)
    c64.set_state(u0)
    dt = dgx.stable_dt(c64, 0.3)
    ctx.rk3_steps(dt, 100)
    c64.rk3_steps(dt, 100)
    a, b = ctx.get_state(), c64.get_state()
    err = np.linalg.norm(a[..., 0] - b[..., 0]) / np.linalg.norm(b[..., 0])
    assert err < 1e-6, err


def test_free_stream_preserved_fp32():
    verts, tris = R.split_quad_mesh(8)
    ctx = dgx.Context(verts, tris, 3, precision="fp32")
    u = np.zeros((ctx.K, ctx.Np, 4))
    # Dubiner phi_0 = 2, so coefficient 0.5 * U is the constant state U
    u[:, 0, :] = 0.5 * np.array([1.0, 0.7, -0.3, 2.9])
    r = ctx.residual(u)
    assert np.abs(r).max() < 1e-5


def test_conservation_sums_and_diagnostics_fp64():
    ora, ctx, u0 = _vortex_setup(12, 2, "dubiner", "fp64")
    ctx.set_state(u0)
    s_gpu = ctx.conservation_sums()
    s_ref = ora.conservation_sums(u0)
    assert np.allclose(s_gpu, s_ref, rtol=1e-13, atol=1e-12)
    assert abs(ctx.max_cfl(0.01) - ora.max_cfl(u0, 0.01)) < 1e-12
    assert np.allclose(ctx.density_range(), ora.density_range(u0), rtol=1e-13)
    assert abs(ctx.kinetic_energy() - ora.kinetic_energy(u0)) < 1e-10
    dt = dgx.stable_dt(ctx, 0.3)
    ctx.rk3_steps(dt, 10)
    s1 = ctx.conservation_sums()
    assert np.abs(s1 - s_gpu).max() < 1e-11 * np.abs(s_gpu).max()


def test_project_and_eval_match_oracle():
    verts, tris = R.split_quad_mesh(7)
    ora = R.Oracle(verts, tris, 3, basis="dubiner")
    ctx = dgx.Context(verts, tris, 3, basis="dubiner", precision="fp64")
    xy = ctx.volume_points()
    xy_ref = ora.volume_points()
    perm = point_permutation(xy, xy_ref)
    ctx.project(R.vortex_state(xy))
    u_gpu = ctx.get_state()
    u_ref = ora.project(R.vortex_state(xy_ref))
    assert _rel(u_gpu, u_ref) < 1e-12
    assert _rel(ctx.eval_volume(), ora.eval_volume(u_ref)[:, perm]) < 1e-12
