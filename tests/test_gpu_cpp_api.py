"""The C++ mirror of the reference API (include/dgfem_amd.hpp) drives the GPU path: run the
compiled vortex driver and compare its diagnostics with the FP64 CPU oracle on the same
mesh, order, dt and step count."""
import os
import subprocess

import numpy as np
import pytest

import dgx
import flows
import refcpu as R

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "discontinuous-galerkin-fem_amd", "lib", "euler_vortex")


def run(*args):
    out = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return {k: float(v) for k, v in (kv.split("=") for kv in out.stdout.split())}


@pytest.mark.parametrize("basis", ["monomial", "dubiner"])
def test_cpp_driver_matches_oracle(basis):
    n, p, steps = 16, 2, 20
    res = run(n, p, steps, "fp64", basis)
    assert int(res["elements"]) == 2 * n * n
    dt = res["dt"]
    # the reference's loop over t = 0, dt, 2 dt, ... < steps dt, frames at step 0 and when the
    # final count is not a multiple of save_every = steps
    assert int(res["steps"]) == dgx.reference_step_count(steps * dt, dt)
    steps = int(res["steps"])
    assert int(res["frames"]) == (1 if steps % 20 == 0 else 3)
    verts, tris = R.split_quad_mesh(n)
    o = R.Oracle(verts, tris, p, basis=basis)
    xy = o.volume_points()
    u0 = o.project(R.vortex_state(xy))
    assert abs(dt - 0.3 / ((2 * p + 1) * o.max_cfl(u0, 1.0))) < 1e-15 * dt * 10
    u, _ = o.rk3(u0, dt, steps)
    s1 = o.conservation_sums(u)
    for v in range(4):
        assert abs(res[f"sum1_{v}"] - s1[v]) < 1e-11 * max(1.0, abs(s1[v]))
        assert abs(res[f"sum1_{v}"] - res[f"sum0_{v}"]) < 1e-11 * max(1.0, abs(s1[v]))
    lo, hi = o.density_range(u)
    assert abs(res["rho_min"] - lo) < 1e-11 and abs(res["rho_max"] - hi) < 1e-11
    Uq = o.eval_volume(u)
    ex = R.vortex_state(xy, t=steps * dt)
    l2 = np.linalg.norm(Uq[..., 0] - ex[..., 0]) / np.linalg.norm(ex[..., 0])
    assert abs(res["l2_rho"] - l2) < 1e-6 * l2 + 1e-12


SHEAR = os.path.join(ROOT, "discontinuous-galerkin-fem_amd", "lib", "ns_shear_flow")


def _shear_conserved(x, y):
    return [1.0, y, 0.0, 1.0 / 0.4 + 0.5 * y * y]


def _shear_oracle_errors(n, points, mesh="tri"):
    # the same run on the FP64 oracle, and the relative L2 error per primitive field
    # (compute_variable_error, T/solver/shear_flow_accuracy_test.cpp:32-79) in numpy
    if mesh == "quad":
        verts, quads = R.rect_quad_mesh(0.0, 1.0, 0.0, 1.0, 1.0 / n)
        ora, u0 = flows.ns_oracle(verts, quads, 2, 5e-6, 1e4, flows.shear_primitive, points)
        u, _ = ora.rk3(u0, 1e-5, 50)
        return flows.l2_errors(ora, u, flows.shear_primitive)
    verts, tris = R.split_quad_mesh(n, 1.0)
    o = R.Oracle(verts, tris, 2, basis="dubiner", periodic=0, physics="ns", mu=5e-6, prandtl=0.72, sigma0=1e4,
                 ns_face_grad="reference", euler_bc_points=points)
    for tag in ("Bottom", "Right", "Top", "Left"):
        o.set_bc(tag, "far_field", fn=_shear_conserved)
    xy = o.volume_points()
    u, _ = o.rk3(o.project(np.array([[_shear_conserved(*q) for q in e] for e in xy])), 1e-5, 50)
    Uq = o.eval_volume(u)
    _, _, w = R.quad_tri(4)
    V = verts[tris]
    det = np.abs((V[:, 1, 0] - V[:, 0, 0]) * (V[:, 2, 1] - V[:, 0, 1]) - (V[:, 2, 0] - V[:, 0, 0]) * (V[:, 1, 1] - V[:, 0, 1]))
    dw = (w[None, :] * det[:, None])[..., None]
    rho = Uq[..., 0]
    W = np.stack([rho, Uq[..., 1] / rho, Uq[..., 2] / rho,
                  0.4 * (Uq[..., 3] - 0.5 * (Uq[..., 1] ** 2 + Uq[..., 2] ** 2) / rho)], axis=-1)
    Wex = np.stack([np.ones_like(rho), xy[..., 1], np.zeros_like(rho), np.ones_like(rho)], axis=-1)
    e2, n2 = ((W - Wex) ** 2 * dw).sum(axis=(0, 1)), (Wex ** 2 * dw).sum(axis=(0, 1))
    return np.where(n2 < 1e-14, np.sqrt(e2), np.sqrt(e2 / np.where(n2 < 1e-14, 1.0, n2)))


@pytest.mark.parametrize("mesh", ["tri", "quad"])
@pytest.mark.parametrize("points", ["reference", "correct"])
def test_cpp_ns_shear_flow_with_function_bc_matches_oracle(points, mesh):
    # the reference's shear-flow accuracy run (far-field condition given as a function of
    # position) through the C++ mirror, on 2 x 20^2 triangles or the test's 20 x 20 quads
    out = subprocess.run([SHEAR, "20", "fp64", points, mesh], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    res = {k: float(v) for k, v in (kv.split("=") for kv in out.stdout.split())}
    assert int(res["elements"]) == (800 if mesh == "tri" else 400)
    # 50 steps with save_every = 50: the reference's loop stores no frame after step 0 (the last
    # one is added only when the step count is not a multiple of save_every), so the reference
    # test's frames.back() is the projected initial state; the driver here reports the error of
    # the context's state after the 50 steps instead
    assert int(res["steps"]) == 50 and int(res["frames"]) == 1
    got = np.array([res[f"err_{v}"] for v in range(4)])
    want = _shear_oracle_errors(20, points, mesh)
    # errors of 1e-8 are differences of O(1) fields: FP64 rounding moves them by ~1e-6 relative
    assert np.all(np.abs(got - want) <= 1e-4 * want + 1e-12), (got, want)
    if points == "correct":
        # the reference test's bounds (:131-134)
        assert got[0] < 5e-4 and got[1] < 5e-3 and got[2] < 5e-4 and got[3] < 5e-3, got
    else:
        # the reference's read of the boundary points perturbs the far-field state enough on
        # triangles that the density error exceeds that test's bound
        assert got[0] > 5e-4, got
