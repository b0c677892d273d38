"""Batched instances: the persistent graph-captured batched step must reproduce the
single-context step of every instance (the same element code; FP64 to 1e-13 relative, FP32 with
the matrix-core orders to 1e-6), for any tuning the autotuner may pick, and odd step counts."""
import numpy as np
import pytest

import dgx
import refcpu as R
import workloads as W

pytestmark = pytest.mark.gpu


def _make(specs, precision):
    meshes = [dgx.split_quad_mesh(s["n"], 10.0) for s in specs]
    b = dgx.Batch(meshes, [s["p"] for s in specs], precision=precision)
    dts = []
    for ctx, s in zip(b.instances, specs):
        xy = ctx.volume_points()
        ctx.project(W.vortex(xy, beta=s["beta"], xc=s["xc"], yc=s["yc"]))
        dts.append(dgx.stable_dt(ctx, 0.3))
    b.set_dt(dts)
    return b, meshes, dts


SPECS = [{"n": 4, "p": 1, "beta": 4.0, "xc": 5.0, "yc": 5.0},
         {"n": 8, "p": 3, "beta": 5.0, "xc": 4.5, "yc": 5.5},
         {"n": 6, "p": 2, "beta": 3.5, "xc": 5.5, "yc": 4.0},
         {"n": 8, "p": 4, "beta": 6.0, "xc": 5.0, "yc": 6.0},
         {"n": 16, "p": 2, "beta": 5.0, "xc": 3.0, "yc": 7.0}]


@pytest.mark.parametrize("layout", ["fused", "per_order", "concurrent"])
@pytest.mark.parametrize("nsteps", [4, 5])
def test_batched_steps_match_single_context(layout, nsteps):
    b, meshes, dts = _make(SPECS, "fp64")
    b.set_tuning(layout, chunk=(1, 2, 4, 1), blocks_per_cu=1)
    u0 = [c.get_state() for c in b.instances]
    b.rk3_steps(nsteps)
    got = [c.get_state() for c in b.instances]
    for (v, t), s, u, g, dt in zip(meshes, SPECS, u0, got, dts):
        ctx = dgx.Context(v, t, s["p"], precision="fp64")
        ctx.set_state(u)
        ctx.rk3_steps(dt, nsteps)
        ref = ctx.get_state()
        assert np.abs(g - ref).max() < 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("layout", ["fused", "per_order", "concurrent"])
@pytest.mark.parametrize("nsteps", [4, 5])
def test_batched_fp32_matches_single_context(layout, nsteps):
    # FP32: every order runs the matrix-core stage in both the batch (tile work items, one
    # launch per order, on one stream or one stream per order) and the single context; fused
    # and per-order launches differ only for per-lane orders (FP64 above).  The single context chooses its own
    # storage offset, so the comparison is to FP32 rounding (1e-6 relative after the steps).
    b, meshes, dts = _make(SPECS, "fp32")
    b.set_tuning(layout, chunk=(1, 2, 4, 1), blocks_per_cu=1)
    u0 = [c.get_state() for c in b.instances]
    b.rk3_steps(nsteps)
    got = [c.get_state() for c in b.instances]
    for (v, t), s, u, g, dt in zip(meshes, SPECS, u0, got, dts):
        ctx = dgx.Context(v, t, s["p"], precision="fp32")
        assert ctx.info.stage_kernel == 1
        ctx.set_state(u)
        ctx.rk3_steps(dt, nsteps)
        ref = ctx.get_state()
        assert np.abs(g - ref).max() < 1e-6 * np.abs(ref).max(), s


def test_batched_instance_matches_oracle():
    b, meshes, dts = _make(SPECS[1:2], "fp64")
    (v, t), s = meshes[0], SPECS[1]
    ora = R.Oracle(v, t, s["p"], basis="dubiner")
    u0 = b.instances[0].get_state()
    b.rk3_steps(6)
    ref, _ = ora.rk3(u0, dts[0], 6)
    assert np.abs(b.instances[0].get_state() - ref).max() < 1e-11 * np.abs(ref).max()


def test_autotune_keeps_results_and_restores_state():
    specs = W.batch_instances(8, seed=3, sizes=(4, 8, 16))
    b, meshes, dts = _make(specs, "fp32")
    before = [c.get_state() for c in b.instances]
    t = b.autotune(2)
    assert t.layout in (0, 1, 2) and all(1 <= c <= 64 for c in t.chunk)
    after = [c.get_state() for c in b.instances]
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    seq = b.profile(4, sequential=True)
    bat = b.profile(4, sequential=False)
    assert seq > 0 and bat > 0
    for x, y in zip(before, [c.get_state() for c in b.instances]):
        assert np.array_equal(x, y)
    b.rk3_steps(3)
    for c in b.instances:
        assert np.isfinite(c.get_state()).all()
