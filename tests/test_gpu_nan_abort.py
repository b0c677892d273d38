"""The reference's NaN abort (run_time_integration, S/solver/compressible_solver_base.cpp:
279-306) on the GPU: the loop stops at the first step whose state holds a NaN, that step's
state is the last one kept, and the frame rule of solve() applies to it.  A vortex stepped at
three times the stable dt goes NaN at step 6 in the FP64 oracle (the wave speed of a trace
with negative pressure is NaN, and the Rusanov max keeps a NaN first operand as std::max
does); a negative-pressure initial state goes NaN at step 1."""
import numpy as np
import pytest

import dgx
import refcpu as R

pytestmark = pytest.mark.gpu


def setup(precision):
    verts, tris = R.split_quad_mesh(8)
    o = R.Oracle(verts, tris, 2, basis="dubiner")
    u0 = o.project(R.vortex_state(o.volume_points()))
    dt = 3 * 0.3 / (5 * o.max_cfl(u0, 1.0))
    ctx = dgx.Context(verts, tris, 2, basis="dubiner", precision=precision)
    return o, ctx, u0, dt


def oracle_first_nan(o, u0, dt, nmax=20):
    u = u0.copy()
    for s in range(1, nmax + 1):
        u, _ = o.rk3(u, dt, 1)
        if np.isnan(u).any():
            return s, u
    return None, u


def test_rk3_run_stops_at_the_oracles_nan_step():
    o, ctx, u0, dt = setup("fp64")
    k, uk = oracle_first_nan(o, u0, dt)
    assert k == 6
    ctx.set_state(u0)
    done, nan = ctx.rk3_run(dt, 10)
    assert nan and done == k
    g = ctx.get_state()
    assert np.array_equal(np.isnan(g), np.isnan(uk))
    fin = np.isfinite(uk)
    assert np.abs(g[fin] - uk[fin]).max() < 1e-9 * np.abs(uk[fin]).max()
    # the step before is clean, and a run that ends before the NaN step reports success
    ctx.set_state(u0)
    assert ctx.rk3_run(dt, k - 1) == (k - 1, False)
    assert not np.isnan(ctx.get_state()).any()
    # the status code of the plain call
    ctx.set_state(u0)
    with pytest.raises(dgx.DGXError) as ei:
        ctx.rk3_steps(dt, 10)
    assert ei.value.status == dgx.DGX_E_NONFINITE
    ctx.close()


def test_rk3_run_nan_state_is_the_replayed_step_fp32():
    # FP32 may go NaN at another step than FP64; whichever it is, the state kept is that
    # step's, bit for bit the state of a plain run of that many steps
    _, ctx, u0, dt = setup("fp32")
    ctx.set_state(u0)
    done, nan = ctx.rk3_run(dt, 12)
    assert nan and 1 <= done <= 12
    g = ctx.get_state()
    ctx.set_state(u0)
    ctx.rk3_steps_async(dt, done)
    ctx.synchronize()
    h = ctx.get_state()
    assert np.array_equal(np.isnan(g), np.isnan(h)) and np.isnan(g).any()
    assert np.array_equal(g[~np.isnan(g)], h[~np.isnan(h)])
    if done > 1:
        ctx.set_state(u0)
        ctx.rk3_steps_async(dt, done - 1)
        ctx.synchronize()
        assert not np.isnan(ctx.get_state()).any()
    ctx.close()


@pytest.mark.parametrize("save_every,frames", [(3, [0, 3]), (4, [0, 4, 6]), (6, [0]), (1, [0, 1, 2, 3, 4, 5])])
def test_solve_keeps_the_reference_frames_up_to_the_nan_step(save_every, frames):
    # the oracle's states at those steps; the NaN step 6 is a frame only when not aligned
    o, ctx, u0, dt = setup("fp64")
    ref = {0: u0}
    u = u0
    for s in range(1, 7):
        u, _ = o.rk3(u, dt, 1)
        ref[s] = u
    solver = dgx.EulerDGSolver(ctx)
    out = solver.solve(lambda xy: R.vortex_state(xy), 10 * dt, dt, save_every)
    assert solver.steps_taken == 6
    assert len(out) == len(frames)
    for f, s in zip(out, frames):
        want = ref[s].reshape(ctx.K, 4 * ctx.Np)
        assert np.array_equal(np.isnan(f), np.isnan(want))
        fin = np.isfinite(want)
        assert np.abs(f[fin] - want[fin]).max() < 1e-9 * np.abs(want[fin]).max()
    ctx.close()


def test_negative_pressure_state_aborts_at_step_one():
    verts, tris = R.split_quad_mesh(4)
    ctx = dgx.Context(verts, tris, 2, basis="dubiner", precision="fp64")
    u = np.zeros((ctx.K, ctx.Np, 4))
    # rho = 1, u = (0.5, 0), p = -0.2: E = p / (gamma - 1) + rho |u|^2 / 2 (phi_0 = 2 in Dubiner)
    u[:, 0, :] = 0.5 * np.array([1.0, 0.5, 0.0, -0.2 / 0.4 + 0.125])
    ctx.set_state(u)
    assert ctx.rk3_run(1e-3, 5) == (1, True)
    assert np.isnan(ctx.get_state()).all()
    ctx.close()
