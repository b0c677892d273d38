"""The C++ mirror of the reference API (include/dgfem_amd.hpp) drives the GPU path: run the
compiled vortex driver and compare its diagnostics with the FP64 CPU oracle on the same
mesh, order, dt and step count."""
import os
import subprocess

import numpy as np
import pytest

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
    assert int(res["frames"]) == 2
    verts, tris = R.split_quad_mesh(n)
    o = R.Oracle(verts, tris, p, basis=basis)
    xy = o.volume_points()
    u0 = o.project(R.vortex_state(xy))
    dt = res["dt"]
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
