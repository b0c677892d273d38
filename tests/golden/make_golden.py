"""Writes the golden fixtures of the CPU-runnable configuration with the FP64 oracle.

Configuration: 32 x 32 periodic split quads on [0,10]^2 (K = 2048), p = 2, monomial basis
(the reference's triangle basis), degree-2p Dunavant volume rule, Gauss-Legendre faces,
isentropic vortex beta = 5 at (5, 5) (rho_inf = u_inf = p_inf = 1, gamma = 1.4), FP64,
dt = 0.3 h_min / ((2p+1) max(|u|+c)) on the projected state.

Fixtures (tests/golden/c1_vortex_p2.npz, host layout [K][Np][4]):
  u0    the L2 projection of the initial condition
  r0    M dU/dt at u0 (assemble_euler_residual)
  u1    one SSP-RK3 step from u0
  u100  100 SSP-RK3 steps from u0
and c1_vortex_p2.json with dt, the SHA-256 of each array's little-endian FP64 bytes and the
relative tolerance a re-computation must meet.

Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import refcpu as R  # noqa: E402

N, P, CFL, STEPS = 32, 2, 0.3, 100
# FP64 restatements agree to reassociation noise; the tolerance leaves room for a different
# (but IEEE) summation order, e.g. another compiler's contraction choices
TOL = {"u0": 1e-13, "r0": 1e-11, "u1": 1e-12, "u100": 1e-10}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def compute():
    verts, tris = R.split_quad_mesh(N, 10.0)
    o = R.Oracle(verts, tris, P, basis="monomial")
    u0 = o.project(R.vortex_state(o.volume_points()))
    dt = CFL / ((2 * P + 1) * o.max_cfl(u0, 1.0))
    r0 = o.residual(u0)
    u1, _ = o.rk3(u0, dt, 1)
    u100, _ = o.rk3(u0, dt, STEPS)
    return dt, {"u0": u0, "r0": r0, "u1": u1, "u100": u100}


def main():
    dt, arrs = compute()
    np.savez_compressed(os.path.join(HERE, "c1_vortex_p2.npz"), **arrs)
    meta = {"config": {"n": N, "elements": 2 * N * N, "order": P, "basis": "monomial", "length": 10.0,
                       "vortex_beta": 5.0, "centre": [5.0, 5.0], "gamma": 1.4, "cfl": CFL, "steps": STEPS},
            "dt": dt, "layout": "[K][Np][4] FP64",
            "arrays": {k: {"sha256": sha(a), "shape": list(a.shape), "rel_tol": TOL[k]} for k, a in arrs.items()},
            "writer": "tests/golden/make_golden.py (FP64 oracle, oracle/refcpu.cpp)"}
    with open(os.path.join(HERE, "c1_vortex_p2.json"), "w") as f:
        json.dump(meta, f, indent=1)
        f.write("\n")
    print(json.dumps({k: v["sha256"][:16] for k, v in meta["arrays"].items()}))


if __name__ == "__main__":
    main()
