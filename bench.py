#!/usr/bin/env python3
"""Benchmark: Euler RHS DOF-stage updates/s on one MI355X.

Workload (default): the scaled isentropic-vortex case -- 256 x 256 periodic split quads on
[0,10]^2 (K = 131072 triangles), p = 3 (Np = 10), Dubiner basis, degree-2p Dunavant volume
rule, vortex beta = 5 at (5,5) plus a seeded i.i.d. +-1% density perturbation per element,
Rusanov flux, SSP-RK3 at dt = 0.3 h_min / ((2p+1) max(|u|+c)).  A "step" is one SSP-RK3
step (three fused stage launches) of the whole mesh; the state is resident in HBM before
the timed region starts.

One DOF-stage update = one conserved-variable modal coefficient advanced by one RK stage:
12 Np K per step.

Output: one JSON line (rank 0).  Besides the contract keys it carries
  roofline      : dominant kernel (the fused FP32 stage kernel), algorithmic bytes per launch
                  / mean launch duration from HIP events on the launch stream, vs 8 TB/s
  cpu_baseline  : the FP64 CPU oracle (port of the reference loop), single thread, timed on
                  a bounded sample of the same mesh
  fp64          : the same workload in FP64 (DOF-stage/s)
  l2_err_fp32_vs_fp64 : per-field relative L2 at volume points after the timed steps
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "discontinuous-galerkin-fem_amd"))

import dgx  # noqa: E402
import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
SEED = 20251203


def vortex_ic(xy, K, Nq, seed):
    pert = None if seed is None else W.density_noise(seed, K, Nq)
    return W.vortex(xy, rho_pert=pert)


def make_ctx(n, p, precision, seed):
    verts, tris = dgx.split_quad_mesh(n, 10.0)
    ctx = dgx.Context(verts, tris, p, basis="dubiner", precision=precision)
    xy = ctx.volume_points()
    ctx.project(vortex_ic(xy, ctx.K, ctx.Nq, seed))
    return ctx, verts, tris, xy


def timed_steps(ctx, dt, steps, warmup):
    u0 = ctx.get_state()
    ctx.rk3_steps_async(dt, warmup)
    ctx.synchronize()
    ctx.set_state(u0)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.rk3_steps_async(dt, steps)
    ctx.synchronize()
    return time.perf_counter() - t0


def l2_rel(a, b):
    return [float(np.linalg.norm(a[..., v] - b[..., v]) / np.linalg.norm(b[..., v])) for v in range(4)]


def cpu_baseline(verts, tris, p, u0, dt, steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refcpu as R
    ora = R.Oracle(verts, tris, p, basis="dubiner")
    _, secs = ora.rk3(u0, dt, 1)  # warm-up step
    _, secs = ora.rk3(u0, dt, steps)
    return 12.0 * ora.Np * ora.K * steps / secs, secs


def load_traffic(workload):
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256, help="mesh: n x n split quads")
    ap.add_argument("--p", type=int, default=3)
    ap.add_argument("--cpu-steps", type=int, default=12, help="oracle steps for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run only a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy_mib:
        calibrate_copy(args.calibrate_copy_mib)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    n, p = args.n, args.p
    workload = f"euler_vortex_n{n}_p{p}"
    ctx32, verts, tris, xy = make_ctx(n, p, "fp32", SEED)
    K, Np = ctx32.K, ctx32.Np
    dt = dgx.stable_dt(ctx32, 0.3)
    u0 = ctx32.get_state()

    secs32 = timed_steps(ctx32, dt, args.steps, args.warmup)
    Uq32 = ctx32.eval_volume()
    dof_stage = 12.0 * Np * K * args.steps
    value = dof_stage / secs32

    # dominant kernel: per-launch HIP events on the context stream
    ctx32.set_state(u0)
    stage_ms, _ = ctx32.rk3_profile(dt, min(args.steps, 50))
    bytes_launch = ctx32.info.bytes_per_stage_alg
    achieved = bytes_launch / (stage_ms * 1e-3) / 1e9

    result = {
        "metric": "Euler RHS DOF-stage updates/s (FP32, L2 err vs FP64 stated)",
        "value": value,
        "unit": "DOF-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs32 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: isentropic vortex + seeded 1% density noise on a periodic split-quad mesh",
        "config": {"workload": workload, "mesh": f"{n}x{n} split quads", "elements": K, "order": p,
                   "modes": Np, "unknowns": K * Np * 4, "basis": "dubiner", "volume_rule": "degree 2p",
                   "dt": dt, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(workload),
                     "kernel": "stage_kernel<P=3,float,Euler>", "launch_ms": stage_ms,
                     "alg_bytes_per_launch": bytes_launch},
    }

    if not args.no_fp64:
        ctx64, _, _, _ = make_ctx(n, p, "fp64", SEED)
        secs64 = timed_steps(ctx64, dt, args.steps, args.warmup)
        Uq64 = ctx64.eval_volume()
        result["fp64"] = {"value": dof_stage / secs64, "ms_per_step": secs64 / args.steps * 1e3}
        err = l2_rel(Uq32, Uq64)
        result["l2_err_fp32_vs_fp64"] = dict(zip(["rho", "rhou", "rhov", "E"], err))
        ctx64.close()

    # analytic oracle on the unperturbed twin (periodic-truncation floor ~1e-5)
    ctxa, _, _, xya = make_ctx(n, p, "fp32", None)
    ctxa.rk3_steps(dt, args.steps)
    Uqa = ctxa.eval_volume()
    exact = vortex_ic_exact(xya, dt * args.steps)
    result["l2_err_fp32_vs_analytic"] = dict(zip(["rho", "rhou", "rhov", "E"], l2_rel(Uqa, exact)))
    ctxa.close()

    if not args.no_cpu:
        v, secs = cpu_baseline(verts, tris, p, u0, dt, args.cpu_steps)
        result["cpu_baseline"] = {"value": v, "unit": "DOF-stage updates/s", "cores": 1, "kind": "port",
                                  "sample": f"{args.cpu_steps} FP64 SSP-RK3 steps of the same {K}-element p={p} mesh "
                                            f"({secs:.1f} s, one thread, oracle/refcpu.cpp)",
                                  "host_cpus": os.cpu_count()}
    ctx32.close()
    print(json.dumps(result))


def calibrate_copy(mib):
    """A plain device-to-device copy of known size: what FETCH_SIZE / WRITE_SIZE read for it
    scales the counters of the stage kernel."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))


def vortex_ic_exact(xy, t):
    return W.vortex(xy, t=t)


if __name__ == "__main__":
    main()
