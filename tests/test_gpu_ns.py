"""GPU parity of the Navier-Stokes residual path (viscous volume/face terms, penalty,
boundary faces) against the FP64 CPU oracle.  FP64 kernels: relative 1e-12 on the
residual; FP32: relative 5e-5."""
import numpy as np
import pytest

import dgx
import refcpu as R

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _tg_setup(n, p, basis, precision, face_grad, mu=6.25e-4):
    verts, tris = R.split_quad_mesh(n, 2 * np.pi)
    ora = R.Oracle(verts, tris, p, basis=basis, physics="ns", mu=mu, prandtl=0.71, sigma0=5.0,
                   ns_face_grad=face_grad)
    wf = dgx.NavierStokesWeakFormulation(1.4, mu, 0.71, 5.0, face_grad=face_grad)
    ctx = dgx.Context(verts, tris, p, weak_form=wf, basis=basis, precision=precision)
    xy = ora.volume_points()
    u0 = ora.project(R.taylor_green_state(xy))
    # add an element-wise perturbation so every gradient and face jump is exercised
    rng = np.random.default_rng(p)
    u0 = u0 + 1e-3 * rng.standard_normal(u0.shape) * np.abs(u0).max(axis=(0, 1))
    return ora, ctx, u0


@pytest.mark.parametrize("face_grad", ["reference", "correct"])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_ns_residual_fp64_matches_oracle(p, face_grad):
    ora, ctx, u0 = _tg_setup(5, p, "dubiner", "fp64", face_grad, mu=0.05)
    r_ref = ora.residual(u0)
    r_gpu = ctx.residual(u0)
    assert _rel(r_gpu, r_ref) < 1e-12


@pytest.mark.parametrize("basis", ["monomial", "dubiner"])
def test_ns_residual_monomial_and_dubiner(basis):
    ora, ctx, u0 = _tg_setup(4, 2, basis, "fp64", "reference", mu=0.02)
    assert _rel(ctx.residual(u0), ora.residual(u0)) < 1e-12


def test_ns_rk3_fp64_matches_oracle():
    ora, ctx, u0 = _tg_setup(8, 3, "dubiner", "fp64", "reference", mu=0.01)
    ctx.set_state(u0)
    dt = dgx.stable_dt(ctx, 0.3)
    u_ref, _ = ora.rk3(u0, dt, 5)
    ctx.rk3_steps(dt, 5)
    assert _rel(ctx.get_state(), u_ref) < 1e-11


def test_ns_residual_fp32_close_to_oracle():
    ora, ctx, u0 = _tg_setup(6, 4, "dubiner", "fp32", "reference", mu=6.25e-4)
    r_ref = ora.residual(u0)
    assert _rel(ctx.residual(u0), r_ref) < 5e-5


def _bc_setup(physics, bc_type, value, precision="fp64"):
    verts, tris = R.split_quad_mesh(4, 1.0)
    kw = dict(physics="ns", mu=0.02, prandtl=0.72, sigma0=5.0) if physics == "ns" else {}
    ora = R.Oracle(verts, tris, 2, basis="dubiner", periodic=0, **kw)
    for tag in ("Bottom", "Right", "Top", "Left"):
        ora.set_bc(tag, bc_type, value)
    wf = dgx.NavierStokesWeakFormulation(1.4, 0.02, 0.72, 5.0) if physics == "ns" else None
    bcs = {tag: (bc_type, value) for tag in ("Bottom", "Right", "Top", "Left")}
    ctx = dgx.Context(verts, tris, 2, weak_form=wf, basis="dubiner", precision=precision, periodic=0, bcs=bcs)
    xy = ora.volume_points()
    u0 = ora.project(R.vortex_state(xy, xc=0.5, yc=0.5, beta=1.0, length=1.0))
    return ora, ctx, u0


FAR = [1.0, 1.0, 0.0, 1 / 0.4 + 0.5]


@pytest.mark.parametrize("physics", ["euler", "ns"])
@pytest.mark.parametrize("bc_type,value", [("far_field", FAR), ("slip_wall", [0, 0, 0, 0]),
                                           ("no_slip_wall", [1.0, 0.0, 0.0, 1.0])])
def test_boundary_faces_match_oracle(physics, bc_type, value):
    ora, ctx, u0 = _bc_setup(physics, bc_type, value)
    assert ctx.info.n_boundary_faces == 16
    assert _rel(ctx.residual(u0), ora.residual(u0)) < 1e-12


def test_boundary_face_without_condition_is_skipped():
    # the assembler skips boundary faces whose condition is unset
    verts, tris = R.split_quad_mesh(3, 1.0)
    ora = R.Oracle(verts, tris, 1, basis="dubiner", periodic=0)
    ctx = dgx.Context(verts, tris, 1, basis="dubiner", precision="fp64", periodic=0)
    xy = ora.volume_points()
    u0 = ora.project(R.vortex_state(xy, xc=0.5, yc=0.5, beta=1.0, length=1.0))
    assert _rel(ctx.residual(u0), ora.residual(u0)) < 1e-12
