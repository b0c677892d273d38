"""The committed golden fixtures of the CPU-runnable configuration (32 x 32 split quads,
p = 2, monomial basis, FP64 isentropic vortex; written by tests/golden/make_golden.py with the
FP64 oracle): the files match their recorded SHA-256, the oracle still reproduces them within
the recorded tolerances, and the fixtures satisfy the invariants of the scheme (global
conservation over 100 steps, a residual whose constant-mode rows sum to zero)."""
import hashlib
import json
import os

import numpy as np
import pytest

import refcpu as R

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load():
    with open(os.path.join(HERE, "c1_vortex_p2.json")) as f:
        meta = json.load(f)
    with np.load(os.path.join(HERE, "c1_vortex_p2.npz"), allow_pickle=False) as z:
        arrs = {k: z[k] for k in z.files}
    return meta, arrs


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_fixtures_match_their_recorded_sha():
    meta, arrs = load()
    assert set(arrs) == set(meta["arrays"])
    for k, a in arrs.items():
        assert list(a.shape) == meta["arrays"][k]["shape"]
        digest = hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
        assert digest == meta["arrays"][k]["sha256"], k


@pytest.fixture(scope="module")
def c1_oracle():
    meta, arrs = load()
    c = meta["config"]
    verts, tris = R.split_quad_mesh(c["n"], c["length"])
    return meta, arrs, R.Oracle(verts, tris, c["order"], basis=c["basis"])


def test_oracle_reproduces_the_fixtures(c1_oracle):
    meta, arrs, o = c1_oracle
    tol = {k: v["rel_tol"] for k, v in meta["arrays"].items()}
    u0 = arrs["u0"]
    assert rel(o.project(R.vortex_state(o.volume_points())), u0) < tol["u0"]
    assert rel(o.residual(u0), arrs["r0"]) < tol["r0"]
    u1, _ = o.rk3(u0, meta["dt"], 1)
    assert rel(u1, arrs["u1"]) < tol["u1"]
    u100, _ = o.rk3(u0, meta["dt"], meta["config"]["steps"])
    assert rel(u100, arrs["u100"]) < tol["u100"]
    # dt is the recorded CFL rule on u0
    p = meta["config"]["order"]
    assert meta["dt"] == pytest.approx(meta["config"]["cfl"] / ((2 * p + 1) * o.max_cfl(u0, 1.0)), rel=1e-14)


def test_fixtures_conserve_and_move(c1_oracle):
    meta, arrs, o = c1_oracle
    s0 = o.conservation_sums(arrs["u0"])
    for k in ("u1", "u100"):
        assert np.abs(o.conservation_sums(arrs[k]) - s0).max() < 1e-12 * np.abs(s0).max()
    r = arrs["r0"]
    assert np.abs(r[:, 0, :].sum(axis=0)).max() < 1e-13 * np.abs(r[:, 0, :]).sum()
    # the vortex has moved: the trajectory is not trivially constant
    assert rel(arrs["u100"], arrs["u0"]) > 1e-2
