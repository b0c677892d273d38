"""The GPU path against the committed golden fixtures of the CPU-runnable configuration
(32 x 32 split quads, p = 2, monomial basis, FP64 isentropic vortex, tests/golden/), through
the C ABI: the residual, one SSP-RK3 step and the configured 100 steps."""
import json
import os

import numpy as np
import pytest

import dgx

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load():
    with open(os.path.join(HERE, "c1_vortex_p2.json")) as f:
        meta = json.load(f)
    with np.load(os.path.join(HERE, "c1_vortex_p2.npz"), allow_pickle=False) as z:
        arrs = {k: z[k] for k in z.files}
    return meta, arrs


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


# FP32 (storage offset, FP32 arithmetic): the residual is a small difference of O(1/h) terms
# (2e-5 of its largest entry); in the monomial basis the inverse mass matrix amplifies FP32
# rounding of the higher modes (1.2e-6 of the largest coefficient after one step)
FP32_TOL = {"r0": 3e-5, "u1": 5e-6, "u100": 5e-5}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gpu_matches_golden_c1(precision):
    # FP64 within the fixtures' own tolerances
    meta, arrs = load()
    c = meta["config"]
    verts, tris = dgx.split_quad_mesh(c["n"], c["length"])
    ctx = dgx.Context(verts, tris, c["order"], basis=c["basis"], precision=precision)
    tol = FP32_TOL if precision == "fp32" else {k: v["rel_tol"] for k, v in meta["arrays"].items()}
    u0 = arrs["u0"]
    errs = {"r0": rel(ctx.residual(u0), arrs["r0"])}
    ctx.set_state(u0)
    ctx.rk3_steps(meta["dt"], 1)
    errs["u1"] = rel(ctx.get_state(), arrs["u1"])
    ctx.set_state(u0)
    ctx.rk3_steps(meta["dt"], c["steps"])
    errs["u100"] = rel(ctx.get_state(), arrs["u100"])
    print(precision, errs)
    assert all(errs[k] < tol[k] for k in errs), (errs, tol)
    ctx.close()
