"""Parity at the benchmark configurations' full sizes.

* Euler vortex, 256 x 256 split quads (K = 131072), p = 3, seeded 1% density noise: one FP64
  SSP-RK3 step against the FP64 CPU oracle (relative 1e-12); the FP32 matrix-core kernel over
  20 steps against the FP64 run (per-field relative L2 1e-5) with the conservation sums kept
  (relative 1e-6, computed in FP64); a uniform state kept uniform (1e-6 of its size).
* Navier-Stokes Taylor-Green, 192 x 192 split quads (K = 73728), p = 4, the reference's
  viscous treatment: the FP64 residual against the oracle (1e-9), and 20 FP32 matrix-core
  steps against the FP64 run (per-field relative L2 at the volume points 1e-5).
* Adjoint, 128 x 128 split quads (K = 32768), p = 3: the FP64 trajectory adjoint against
  central differences of the oracle's trajectory along random directions (dot-product test,
  relative 1e-6), and the FP32 matrix-core adjoint against the FP64 one (relative 1e-4).
* Batch, the 64 instances of the batched benchmark (N in {16..128}, p in 1..4): every instance
  after three batched steps against its own single-context run (FP32, relative 1e-6).
"""
import numpy as np
import pytest

import dgx
import refcpu as R
import workloads as W

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _l2(a, b):
    return [float(np.linalg.norm(a[..., v] - b[..., v]) / np.linalg.norm(b[..., v])) for v in range(4)]


def _vortex_c2(ctx):
    xy = ctx.volume_points()
    ctx.project(W.vortex(xy, rho_pert=W.density_noise(20251203, ctx.K, ctx.Nq)))
    return ctx.get_state()


@pytest.fixture(scope="module")
def c2_mesh():
    return dgx.split_quad_mesh(256, 10.0)


def test_c2_fp64_step_matches_oracle(c2_mesh):
    verts, tris = c2_mesh
    ctx = dgx.Context(verts, tris, 3, basis="dubiner", precision="fp64")
    u0 = _vortex_c2(ctx)
    dt = dgx.stable_dt(ctx, 0.3)
    ctx.rk3_steps(dt, 1)
    ora = R.Oracle(verts, tris, 3, basis="dubiner")
    u1, _ = ora.rk3(u0, dt, 1)
    assert _rel(ctx.get_state(), u1) < 1e-12


def test_c2_fp32_matrix_core_against_fp64_and_conservation(c2_mesh):
    verts, tris = c2_mesh
    c32 = dgx.Context(verts, tris, 3, basis="dubiner", precision="fp32")
    assert c32.info.stage_kernel == 1  # the matrix-core tile kernel
    c64 = dgx.Context(verts, tris, 3, basis="dubiner", precision="fp64")
    u0 = _vortex_c2(c64)
    c32.set_state(u0)
    dt = dgx.stable_dt(c64, 0.3)
    s0 = np.array(c64.conservation_sums())
    c32.rk3_steps(dt, 20)
    c64.rk3_steps(dt, 20)
    assert max(_l2(c32.get_state(), c64.get_state())) < 1e-5
    s32 = np.array(c32.conservation_sums())
    assert np.all(np.abs(s32 - s0) <= 1e-6 * np.abs(s0).max())


def test_c2_fp32_free_stream_preserved(c2_mesh):
    verts, tris = c2_mesh
    ctx = dgx.Context(verts, tris, 3, basis="dubiner", precision="fp32")
    W0 = np.array([1.0, 1.0, 0.5, 1.0 / 0.4 + 0.5 * 1.25])
    xy = ctx.volume_points()
    ctx.project(np.broadcast_to(W0, xy.shape[:-1] + (4,)).copy())
    u0 = ctx.get_state()
    ctx.rk3_steps(dgx.stable_dt(ctx, 0.3), 10)
    assert np.abs(ctx.get_state() - u0).max() < 1e-6 * np.abs(W0).max()


def test_c3_ns_residual_matches_oracle():
    verts, tris = dgx.split_quad_mesh(192, 2 * np.pi)
    wf = dgx.NavierStokesWeakFormulation(1.4, 1.0 / 1600.0, 0.71, 5.0, face_grad="reference")
    c64 = dgx.Context(verts, tris, 4, weak_form=wf, basis="dubiner", precision="fp64")
    c64.project(W.taylor_green(c64.volume_points()))
    u0 = c64.get_state()
    ora = R.Oracle(verts, tris, 4, basis="dubiner", physics="ns", mu=1.0 / 1600.0, prandtl=0.71)
    r_ref = ora.residual(u0)
    # FP64 summation order differs from the oracle's (element-centric faces, tables folded into
    # the products) and the residual here is a small difference of viscous face and penalty
    # terms that grow as 1/h: 6e-11 of its largest entry measured, held to 1e-9
    assert _rel(c64.residual(u0), r_ref) < 1e-9

def test_c3_ns_fp32_steps_against_fp64():
    # At Mach 0.1 the pressure is 1 / (gamma M^2) = 71 times the dynamic pressure, and the
    # residual of this divergence-free initial state is a small difference of pressure fluxes:
    # FP32 rounding of those fluxes is 3% of the residual's largest entry here (the per-lane and
    # matrix-core FP32 kernels alike), falling to 1e-2 at 24 x 24 (the residual shrinks as
    # h^2, the rounding as h).  The state stays close: 20 steps at the benchmark's dt, per-field
    # relative L2 against the FP64 run held to 1e-5.
    verts, tris = dgx.split_quad_mesh(192, 2 * np.pi)
    wf = dgx.NavierStokesWeakFormulation(1.4, 1.0 / 1600.0, 0.71, 5.0, face_grad="reference")
    c64 = dgx.Context(verts, tris, 4, weak_form=wf, basis="dubiner", precision="fp64")
    c64.project(W.taylor_green(c64.volume_points()))
    c32 = dgx.Context(verts, tris, 4, weak_form=wf, basis="dubiner", precision="fp32")
    assert c32.info.stage_kernel == 1  # the matrix-core tile kernel
    c32.set_state(c64.get_state())
    dt = dgx.stable_dt(c64, 0.05)
    c32.rk3_steps(dt, 20)
    c64.rk3_steps(dt, 20)
    assert max(_l2(c32.eval_volume(), c64.eval_volume())) < 1e-5


def test_c4_adjoint_full_size():
    verts, tris = dgx.split_quad_mesh(128, 10.0)
    c64 = dgx.Context(verts, tris, 3, basis="dubiner", precision="fp64")
    xy = c64.volume_points()
    # a transverse velocity keeps the face points off the u.n = 0 kink of the Rusanov flux
    U = W.vortex(xy, xc=4.71, yc=5.29, rho_pert=W.density_noise(11, c64.K, c64.Nq))
    U[..., 2] += 0.3 * U[..., 0]
    U[..., 3] += 0.5 * 0.09 * U[..., 0] + 0.3 * (U[..., 2] - 0.3 * U[..., 0])
    c64.project(U)
    u0 = c64.get_state()
    dt = dgx.stable_dt(c64, 0.3)
    nsteps = 3
    lam_T = np.random.default_rng(5).standard_normal(u0.shape)
    lam0 = c64.rk3_vjp(dt, nsteps, 2, lam_T)
    ora = R.Oracle(verts, tris, 3, basis="dubiner")
    # a central difference whose stencil straddles a kink of the Rusanov speed max(s_L, s_R)
    # returns the mean of the one-sided slopes: a direction passes at any of three steps
    for d in [np.random.default_rng(s).standard_normal(u0.shape) for s in (21, 22)]:
        gd = float(np.sum(lam0 * d))
        errs = []
        for h in (1e-6, 1e-7, 1e-8):
            fd = float(np.sum((ora.rk3(u0 + h * d, dt, nsteps)[0] - ora.rk3(u0 - h * d, dt, nsteps)[0])
                              / (2 * h) * lam_T))
            errs.append(abs(gd - fd) / abs(fd))
            if errs[-1] < (1e-6 if h > 1e-8 else 3e-6):
                break
        else:
            raise AssertionError(errs)
    c32 = dgx.Context(verts, tris, 3, basis="dubiner", precision="fp32")
    c32.set_state(u0)
    g32 = c32.rk3_vjp(dt, nsteps, 2, lam_T)
    assert np.linalg.norm(g32 - lam0) / np.linalg.norm(lam0) < 1e-4


def test_c5_full_batch_matches_single_contexts():
    specs = W.batch_instances(64, seed=5)
    meshes = [dgx.split_quad_mesh(s["n"], 10.0) for s in specs]
    b = dgx.Batch(meshes, [s["p"] for s in specs], precision="fp32")
    dts = []
    for ctx, s in zip(b.instances, specs):
        ctx.project(W.vortex(ctx.volume_points(), beta=s["beta"], xc=s["xc"], yc=s["yc"]))
        dts.append(dgx.stable_dt(ctx, 0.3))
    b.set_dt(dts)
    u0 = [c.get_state() for c in b.instances]
    b.rk3_steps(3)
    got = [c.get_state() for c in b.instances]
    for (v, t), s, u, g, dt in zip(meshes, specs, u0, got, dts):
        ctx = dgx.Context(v, t, s["p"], precision="fp32")
        ctx.set_state(u)
        ctx.rk3_steps(dt, 3)
        assert _rel(g, ctx.get_state()) < 1e-6, s
        ctx.close()
    b.close()
