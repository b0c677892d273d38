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
  compute       : the compute side for the same kernel: SIMD issue cycles (VALU + FP32 MFMA,
                  which serialise on gfx950) from the committed SQ counters over the live
                  launch time, and the FP32 matrix-product rate vs 157.3 TFLOP/s
  fp64          : the same workload in FP64 (DOF-stage/s)
  accuracy      : per-field relative L2 at the volume points after the workload's 200 steps
                  (whatever --steps is): FP32 vs the FP64 run, and vs the analytic vortex on
                  the unperturbed twin
  bf16_volume_variant : the same run with bfloat16 operands in the volume products (FP32
                  accumulation and state): DOF-stage/s and per-field L2 vs FP64 at 200 steps
  adjoint       : the metric's second half at the adjoint workload (K = 32768, 200 steps,
                  checkpoints every 20): reverse sweep / forward and whole gradient / forward,
                  medians of three profiled runs
  cpu_baseline  : the reference algorithm (oracle/refcpu.cpp) built with the reference's
                  Release flags on this host, FP64, median of 3 runs on OMP_NUM_THREADS
                  threads (value) and on one thread, with and without -ffast-math; lscpu model
  fp64_vs_oracle_full_size : max relative difference of the GPU FP64 state and the CPU
                  baseline's single-thread state after the same steps at the full size
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "discontinuous-galerkin-fem_amd"))

import dgx  # noqa: E402
import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
SEED = 20251203
ACC_STEPS = 200   # the headline workload's step count: its accuracy condition is stated there
ADJ_STEPS = 200   # the adjoint workload's step count


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


def host_cpu():
    """(lscpu model name, nproc) of the host running the bench."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=20).stdout.strip())
    except (OSError, subprocess.SubprocessError, ValueError):
        nproc = os.cpu_count()
    return model, nproc


def cpu_baseline(verts, tris, p, u0, dt, steps, threads, runs=3):
    """The reference algorithm on the host's CPU: the oracle's loop structure (oracle/refcpu.cpp)
    compiled as the reference's Release build (-O3 -march=native, with and without -ffast-math;
    built here, on the timing host, by `make -B -C oracle timing`), one warm-up step, then the
    median of `runs` timed runs of `steps` steps on one thread (the reference is
    single-threaded) and of 4 x `steps` steps on `threads` OpenMP threads.  Also returns the
    state of the single-thread run of `steps` steps from u0 (the full-size FP64 parity check).
    Returns ({variant: {...}}, u_steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refcpu as R
    subprocess.run(["make", "-B", "-s", "-C", os.path.join(ROOT, "oracle"), "timing"], check=True,
                   stdout=subprocess.DEVNULL)
    out, u_steps = {}, None
    for variant in ("native", "fastmath"):
        ora = R.Oracle(verts, tris, p, basis="dubiner", build=variant)
        dof = 12.0 * ora.Np * ora.K
        ora.rk3(u0, dt, 1)
        one = []
        for _ in range(runs):
            u, secs = ora.rk3(u0, dt, steps)
            one.append(secs)
        if variant == "native":
            u_steps = u
        ora.rk3_threads(u0, dt, 1, threads)
        many = [ora.rk3_threads(u0, dt, 4 * steps, threads)[1] for _ in range(runs)]
        s1, sn = float(np.median(one)), float(np.median(many))
        out[variant] = {"single_thread": dof * steps / s1, "single_thread_s": one,
                        "all_threads": dof * 4 * steps / sn, "all_threads_s": many}
        del ora
    return out, u_steps


def load_counters(workload, kernel=None):
    """Per-launch counter summary (HBM bytes, instruction counts) of `kernel` committed under
    profiles/, or None when it is missing or was measured on another kernel."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if kernel is not None and t.get("kernel") != kernel:
        return None
    return t


def load_traffic(workload, kernel=None):
    """HBM bytes per launch of `kernel` from the committed counter summary, or None."""
    t = load_counters(workload, kernel)
    return t.get("bytes_per_launch") if t else None


CLOCK_HZ = 2.4e9           # MI355X peak engine clock
FP32_MFMA_PEAK_TFLOPS = 157.3


def compute_fraction(workload, kernel, launch_ms):
    """Compute side of the roofline from the committed SQ counters and the live launch time.
    On gfx950 the FP32 matrix products and the VALU share one issue path (they serialise), so
    the compute bound is the SIMD issue time: VALU instructions x 4 cycles plus the measured
    FP32-MFMA busy cycles, per SIMD, over the launch time at the peak clock."""
    t = load_counters(workload, kernel)
    if not t or "valu_insts" not in t:
        return None
    simds = t["simds"]
    cycles = (4.0 * t["valu_insts"] + t["mfma_busy_cycles"]) / simds
    launch_s = launch_ms * 1e-3
    mfma_tflops = t["mfma_insts"] * 2048.0 / launch_s / 1e12
    return {"bound": "simd issue (VALU + FP32 MFMA)", "issue_cycles_per_simd": cycles,
            "issue_floor_ms": cycles / CLOCK_HZ * 1e3, "frac": cycles / (launch_s * CLOCK_HZ),
            "mfma_tflops": mfma_tflops, "mfma_peak_tflops": FP32_MFMA_PEAK_TFLOPS,
            "mfma_frac": mfma_tflops / FP32_MFMA_PEAK_TFLOPS}


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
    ap.add_argument("--physics", default="euler", choices=["euler", "ns"],
                    help="quad workload only: Euler, or Navier-Stokes (mu = 1e-3, Pr = 0.71, the consistent "
                         "viscous scheme, dt capped by the explicit viscous limit)")
    ap.add_argument("--workload", default="euler",
                    choices=["euler", "ns", "quad", "adjoint", "batch"],
                    help="euler: the headline vortex case; ns: Taylor-Green Navier-Stokes; quad: the vortex "
                         "on quadrilaterals; adjoint: gradient of the final kinetic energy; batch: 64 mixed instances")
    ap.add_argument("--no-autotune", action="store_true", help="batch: time the default tuning only")
    args = ap.parse_args()
    if args.calibrate_copy_mib:
        calibrate_copy(args.calibrate_copy_mib)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if args.workload != "euler":
        {"ns": bench_ns, "quad": bench_quad, "adjoint": bench_adjoint, "batch": bench_batch}[args.workload](args)
        return

    n, p = args.n, args.p
    workload = f"euler_vortex_n{n}_p{p}"
    ctx32, verts, tris, xy = make_ctx(n, p, "fp32", SEED)
    K, Np = ctx32.K, ctx32.Np
    dt = dgx.stable_dt(ctx32, 0.3)
    u0 = ctx32.get_state()

    secs32 = timed_steps(ctx32, dt, args.steps, args.warmup)
    dof_stage = 12.0 * Np * K * args.steps
    value = dof_stage / secs32

    # dominant kernel: per-launch HIP events on the context stream
    ctx32.set_state(u0)
    stage_ms, _ = ctx32.rk3_profile(dt, min(args.steps, 50))
    bytes_launch = ctx32.info.bytes_per_stage_alg
    achieved = bytes_launch / (stage_ms * 1e-3) / 1e9
    kname = ("stage_mfma_kernel_p%d" % p) if ctx32.info.stage_kernel == 1 else "stage_kernel<%d,float,Euler>" % p

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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(workload, kname),
                     "kernel": kname, "launch_ms": stage_ms,
                     "alg_bytes_per_launch": bytes_launch},
        "compute": compute_fraction(workload, kname, stage_ms),
    }

    FIELDS = ["rho", "rhou", "rhov", "E"]
    ctx64 = None
    if not args.no_fp64:
        ctx64, _, _, _ = make_ctx(n, p, "fp64", SEED)
        secs64 = timed_steps(ctx64, dt, args.steps, args.warmup)
        result["fp64"] = {"value": dof_stage / secs64, "ms_per_step": secs64 / args.steps * 1e3}
        # the workload's accuracy condition at its own 200 steps, whatever --steps is: FP32
        # (and the bfloat16 volume variant) against the FP64 run from the same state and dt
        acc = {"steps": ACC_STEPS}
        ctx64.set_state(u0)
        ctx64.rk3_steps(dt, ACC_STEPS)
        Uq64 = ctx64.eval_volume()
        ctx32.set_state(u0)
        ctx32.rk3_steps(dt, ACC_STEPS)
        acc["l2_err_fp32_vs_fp64"] = dict(zip(FIELDS, l2_rel(ctx32.eval_volume(), Uq64)))
        if p in (3, 4):
            ctxb, _, _, _ = make_ctx(n, p, "fp32_bf16vol", SEED)
            secsb = timed_steps(ctxb, dt, args.steps, args.warmup)
            ctxb.set_state(u0)
            stage_b, _ = ctxb.rk3_profile(dt, min(args.steps, 50))
            ctxb.set_state(u0)
            ctxb.rk3_steps(dt, ACC_STEPS)
            result["bf16_volume_variant"] = {
                "value": dof_stage / secsb, "ms_per_step": secsb / args.steps * 1e3, "launch_ms": stage_b,
                "kernel": "stage_mfma_bf16_kernel_p%d" % p,
                "l2_err_vs_fp64_200_steps": dict(zip(FIELDS, l2_rel(ctxb.eval_volume(), Uq64)))}
            ctxb.close()
        # analytic oracle on the unperturbed twin (periodic-truncation floor ~1e-5)
        ctxa, _, _, xya = make_ctx(n, p, "fp32", None)
        ctxa.rk3_steps(dt, ACC_STEPS)
        acc["l2_err_fp32_vs_analytic"] = dict(zip(FIELDS, l2_rel(ctxa.eval_volume(),
                                                                 vortex_ic_exact(xya, dt * ACC_STEPS))))
        ctxa.close()
        result["accuracy"] = acc

    # the other half of the metric: adjoint / forward time at the adjoint workload (128 x 128
    # split quads, K = 32768, p = 3, FP32, 200 steps, checkpoints every 20), whatever --steps is
    result["adjoint"] = adjoint_ratios(128, p, ADJ_STEPS, 20)
    # many small instances at once: the 64-instance batch at its default tuning, 100 steps
    result["batch"] = batch_rates(100, autotune=False)
    result["batch"]["over_headline_value"] = result["batch"]["value"] / value

    if not args.no_cpu:
        # the box's CPU share is what OMP_NUM_THREADS says (not os.cpu_count())
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        model, nproc = host_cpu()
        var, u_cpu = cpu_baseline(verts, tris, p, u0, dt, args.cpu_steps, threads)
        fm = var["fastmath"]
        result["cpu_baseline"] = {
            "value": fm["all_threads"], "unit": "DOF-stage updates/s", "cores": threads, "kind": "port",
            "sample": f"median of 3 runs of {4 * args.cpu_steps} FP64 SSP-RK3 steps of the same {K}-element p={p} "
                      f"mesh from the same state and dt on {threads} OpenMP threads, reference algorithm restated "
                      f"(oracle/refcpu.cpp) and built as the reference's Release build (-O3 -march=native "
                      f"-ffast-math); single thread: median of 3 runs of {args.cpu_steps} steps",
            "single_thread_value": fm["single_thread"],
            "variants": {"O3_march_native_ffast_math": fm, "O3_march_native": var["native"]},
            "cpu_model": model, "nproc": nproc}
        if ctx64 is not None:
            # full-size FP64 parity: the GPU's FP64 state after the same steps from the same state
            ctx64.set_state(u0)
            ctx64.rk3_steps(dt, args.cpu_steps)
            g = ctx64.get_state()
            result["fp64_vs_oracle_full_size"] = {
                "steps": args.cpu_steps, "elements": K,
                "max_rel_diff": float(np.abs(g - u_cpu).max() / np.abs(u_cpu).max())}
    if ctx64 is not None:
        ctx64.close()
    ctx32.close()
    print(json.dumps(result))


def adjoint_ratios(n, p, steps, every):
    """Reverse sweep and whole gradient (checkpointed forward + reverse) over a plain forward
    run of the same steps, gradient of the final kinetic energy w.r.t. the initial state."""
    ctx, _, _, _ = make_ctx(n, p, "fp32", SEED)
    dt = dgx.stable_dt(ctx, 0.3)
    ctx.rk3_vjp_profile(dt, 3 * every, every)  # warm-up: side stream, code paths
    runs = [ctx.rk3_vjp_profile(dt, steps, every) for _ in range(3)]
    rev_ratio = [r / f for f, _, r in runs]
    grad_ratio = [(a + r) / f for f, a, r in runs]
    fwd, ckpt, rev = (float(np.median([x[k] for x in runs])) for k in range(3))
    out = {"workload": f"euler_vortex_adjoint_n{n}_p{p}", "elements": ctx.K, "steps": steps,
           "checkpoint_every": every, "runs": 3, "forward_ms": fwd, "checkpointed_forward_ms": ckpt,
           "reverse_sweep_ms": rev, "reverse_over_forward": float(np.median(rev_ratio)),
           "gradient_over_forward": float(np.median(grad_ratio)),
           "gradient_over_forward_runs": grad_ratio}
    ctx.close()
    return out


def _line(metric, value, unit, steps, warmup, ms_per_step, dtype, config, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": dtype, "data": "synthetic", "config": config}
    out.update(extra)
    print(json.dumps(out))


def bench_ns(args):
    """Taylor-Green vortex, compressible Navier-Stokes, 192 x 192 split quads on [0, 2 pi]^2
    (K = 73728), p = 4, Mach 0.1, Re 1600, Pr 0.71, FP32, reference face-gradient read."""
    n = 192 if args.n == 256 else args.n
    p = 4 if args.p == 3 else args.p
    steps = args.steps if args.steps != 200 else 500
    verts, tris = dgx.split_quad_mesh(n, 2 * np.pi)
    wf = dgx.NavierStokesWeakFormulation(1.4, 1.0 / 1600.0, 0.71, 5.0, face_grad="reference")
    ctx = dgx.Context(verts, tris, p, weak_form=wf, basis="dubiner", precision="fp32")
    ctx.project(W.taylor_green(ctx.volume_points()))
    # The reference's viscous treatment (its heat flux runs up the temperature gradient) lets
    # this flow grow without bound from t ~ 0.015 whatever the step; CFL 0.05 keeps the 500
    # steps (t = 0.008) inside the bounded window.  The work per step does not depend on dt.
    dt = dgx.stable_dt(ctx, 0.05)
    secs = timed_steps(ctx, dt, steps, args.warmup)
    Uq32 = ctx.eval_volume()
    extra = {}
    if not args.no_fp64:
        # the same run in FP64 (per-lane kernel): per-field relative L2 at the volume points
        c64 = dgx.Context(verts, tris, p, weak_form=wf, basis="dubiner", precision="fp64")
        c64.project(W.taylor_green(c64.volume_points()))
        c64.rk3_steps(dt, steps)
        extra["l2_err_fp32_vs_fp64"] = dict(zip(["rho", "rhou", "rhov", "E"], l2_rel(Uq32, c64.eval_volume())))
    ctx.set_state(ctx.get_state())
    stage_ms, _ = ctx.rk3_profile(dt, 20)
    K, Np = ctx.K, ctx.Np
    bytes_launch = ctx.info.bytes_per_stage_alg
    # one RK stage = one launch of the matrix-core tile kernel
    kname = ("ns_tile_quirk_kernel_p%d" % p if ctx.info.stage_kernel == 1 else "stage_kernel<%d,float,NS>" % p)
    _line("NS RHS DOF-stage updates/s (FP32)", 12.0 * Np * K * steps / secs, "DOF-stage updates/s", steps,
          args.warmup, secs / steps * 1e3, "f32",
          {"workload": f"ns_taylor_green_n{n}_p{p}", "elements": K, "order": p, "mu": 1 / 1600, "dt": dt},
          roofline={"bound": "hbm", "achieved": bytes_launch / (stage_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": bytes_launch / (stage_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "traffic": load_traffic(f"ns_taylor_green_n{n}_p{p}", kname), "kernel": kname,
                    "launch_ms": stage_ms, "alg_bytes_per_launch": bytes_launch},
          # NS p = 4 sits near the ridge: the compute side beside the bandwidth side
          compute=compute_fraction(f"ns_taylor_green_n{n}_p{p}", kname, stage_ms), **extra)


def bench_quad(args):
    """The isentropic vortex on quadrilaterals: 256 x 256 periodic structured quads on [0,10]^2
    (K = 65536), p = 3 (Legendre tensor basis, Np = 16, 4 x 4 Gauss-Legendre volume rule), FP32,
    seeded 1% density noise, dt at CFL 0.3, 200 SSP-RK3 steps; the FP64 run beside it and the
    FP64 CPU oracle on a bounded sample of the same mesh."""
    n = args.n
    p = args.p
    ns = args.physics == "ns"
    verts, quads = dgx.rect_quad_mesh(0.0, 10.0, 0.0, 10.0, 10.0 / n)
    # Navier-Stokes: the consistent viscous scheme (true face gradients, +{Fv}.n, heat flux as
    # conduction; the reference's treatment grows without bound here), and dt capped by the
    # explicit viscous limit: at h = 0.039 the convective CFL 0.3 step exceeds it (the FP64
    # oracle reaches non-finite values within 20 steps there)
    wf = (dgx.NavierStokesWeakFormulation(1.4, 1e-3, 0.71, 5.0, face_grad="correct", viscous_average="consistent",
                                          heat_flux="conduction") if ns else None)

    def make(precision):
        c = dgx.Context(verts, quads, p, weak_form=wf, precision=precision)
        xy = c.volume_points()
        c.project(vortex_ic(xy, c.K, c.Nq, SEED))
        return c

    ctx = make("fp32")
    K, Np = ctx.K, ctx.Np
    dt = dgx.stable_dt(ctx, 0.3)
    if ns:
        # explicit viscous limit ~ h^2 / (mu (p+1)^4) (penalty and face gradients scale with
        # (p+1)^2 / h each); the factor 0.02 keeps the n = 256 run finite (the FP64 oracle goes
        # non-finite at the convective CFL 0.3 step there) and leaves coarser meshes at CFL 0.3
        h = 10.0 / n
        dt = min(dt, 0.02 * h * h / (1e-3 * (p + 1) ** 4))
    u0 = ctx.get_state()
    secs = timed_steps(ctx, dt, args.steps, args.warmup)
    lo, hi = ctx.density_range()
    if not (np.isfinite(lo) and np.isfinite(hi) and lo > 0):
        print(f"bench.py: the stepped state is not physical (density range [{lo}, {hi}]); no line printed",
              file=sys.stderr)
        sys.exit(3)
    Uq32 = ctx.eval_volume()
    ctx.set_state(u0)
    stage_ms, _ = ctx.rk3_profile(dt, min(args.steps, 50))
    bytes_launch = ctx.info.bytes_per_stage_alg
    workload = f"{'ns' if ns else 'euler'}_vortex_quads_n{n}_p{p}"
    kname = ("quad_tensor_stage_kernel<%d,float>" % p if ctx.info.stage_kernel == 2
             else "quad_stage_kernel<%d,float,%s>" % (p, "NS" if ns else "Euler"))
    extra = {}
    if not args.no_fp64:
        c64 = make("fp64")
        secs64 = timed_steps(c64, dt, args.steps, args.warmup)
        extra["fp64"] = {"value": 12.0 * Np * K * args.steps / secs64, "ms_per_step": secs64 / args.steps * 1e3}
        extra["l2_err_fp32_vs_fp64"] = dict(zip(["rho", "rhou", "rhov", "E"], l2_rel(Uq32, c64.eval_volume())))
        c64.close()
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import refcpu as R
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ora = (R.Oracle(verts, quads, p, physics="ns", mu=1e-3, prandtl=0.71, sigma0=5.0, ns_face_grad="correct",
                        ns_average="consistent", heat_flux="conduction") if ns else R.Oracle(verts, quads, p))
        ora.rk3_threads(u0, dt, 1, threads)
        _, s = ora.rk3_threads(u0, dt, 4 * args.cpu_steps, threads)
        extra["cpu_baseline"] = {"value": 12.0 * Np * K * 4 * args.cpu_steps / s, "unit": "DOF-stage updates/s",
                                 "cores": threads, "kind": "port",
                                 "sample": f"{4 * args.cpu_steps} FP64 SSP-RK3 steps of the same {K}-quad p={p} mesh "
                                           f"on {threads} OpenMP threads ({s:.1f} s, oracle/refcpu.cpp)"}
    ctx.close()
    achieved = bytes_launch / (stage_ms * 1e-3) / 1e9
    _line(("NS" if ns else "Euler") + " RHS DOF-stage updates/s on quadrilaterals (FP32)", 12.0 * Np * K * args.steps / secs,
          "DOF-stage updates/s", args.steps, args.warmup, secs / args.steps * 1e3, "f32",
          {"workload": workload, "mesh": f"{n}x{n} structured quads", "elements": K, "order": p, "modes": Np,
           "basis": "legendre", "volume_rule": f"{p + 1}x{p + 1} Gauss-Legendre", "dt": dt},
          roofline={"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(workload, kname), "kernel": kname,
                    "launch_ms": stage_ms, "alg_bytes_per_launch": bytes_launch},
          compute=compute_fraction(workload, kname, stage_ms), **extra)


def bench_adjoint(args):
    """Gradient of J = int KE(t_final) w.r.t. the initial coefficients: 128 x 128 split quads
    (K = 32768), p = 3, FP32, 200 SSP-RK3 steps, checkpoints every 20, reverse sweep with
    recompute.  Reports adjoint time / forward time."""
    n = 128 if args.n == 256 else args.n
    steps = args.steps
    ctx, _, _, _ = make_ctx(n, args.p, "fp32", SEED)
    dt = dgx.stable_dt(ctx, 0.3)
    fwd_ms, ckpt_ms, rev_ms = ctx.rk3_vjp_profile(dt, steps, 20)
    K, Np = ctx.K, ctx.Np
    # the adjoint is the reverse sweep (per-segment recompute from the checkpoints + stage
    # adjoints); the whole gradient also runs the checkpointed forward sweep
    _line("adjoint sweep time / forward time (Euler, FP32)", rev_ms / fwd_ms, "ratio", steps, 1, rev_ms / steps, "f32",
          {"workload": f"euler_vortex_adjoint_n{n}_p{args.p}", "elements": K, "order": args.p,
           "checkpoint_every": 20, "dt": dt},
          higher_is_better=False, forward_ms=fwd_ms, checkpointed_forward_ms=ckpt_ms, reverse_sweep_ms=rev_ms,
          gradient_total_ratio=(ckpt_ms + rev_ms) / fwd_ms,
          forward_dof_stage_per_s=12.0 * Np * K * steps / (fwd_ms * 1e-3))


def make_batch64():
    """64 independent vortex instances (N in {16,32,64,128}, p in {1..4}, beta in [3,6], centre
    in [3,7]^2, drawn from splitmix64), FP32, each at its own stable dt."""
    specs = W.batch_instances(64, seed=5)
    meshes = [dgx.split_quad_mesh(s["n"], 10.0) for s in specs]
    b = dgx.Batch(meshes, [s["p"] for s in specs], precision="fp32")
    dts = []
    for ctx, s in zip(b.instances, specs):
        ctx.project(W.vortex(ctx.volume_points(), beta=s["beta"], xc=s["xc"], yc=s["yc"]))
        dts.append(dgx.stable_dt(ctx, 0.3))
    b.set_dt(dts)
    return b


def batch_rates(steps, autotune):
    """Batched graph-captured steps (default tuning, then the autotuner's if asked) vs the same
    instances run one at a time with the single-context launches."""
    b = make_batch64()
    dof = sum(12.0 * c.Np * c.K for c in b.instances)
    out = {"workload": "euler_vortex_batch64", "instances": 64, "elements_total": sum(c.K for c in b.instances),
           "steps": steps}
    ms = b.profile(steps)
    out["default_tuning"] = b.tuning().as_dict()
    if autotune:
        out["untuned_value"] = dof / (ms * 1e-3)
        out["tuning"] = b.autotune(4).as_dict()
        ms = b.profile(steps)
    seq_ms = b.profile(steps, sequential=True)
    out.update({"value": dof / (ms * 1e-3), "ms_per_step": ms, "sequential_value": dof / (seq_ms * 1e-3),
                "sequential_ms_per_step": seq_ms, "speedup_vs_sequential": seq_ms / ms})
    b.close()
    return out


def bench_batch(args):
    """The 64-instance batch: 100 steps (or --steps), autotuned (unless --no-autotune), vs one
    instance at a time."""
    r = batch_rates(100 if args.steps == 200 else args.steps, autotune=not args.no_autotune)
    _line("batched Euler DOF-stage updates/s (FP32, 64 instances)", r.pop("value"), "DOF-stage updates/s",
          r.pop("steps"), 2, r.pop("ms_per_step"), "f32",
          {"workload": r.pop("workload"), "instances": r.pop("instances"), "elements_total": r.pop("elements_total")},
          **r)


def calibrate_copy(mib):
    """A plain device-to-device copy of known size: what FETCH_SIZE / WRITE_SIZE read for it
    scales the counters of the stage kernel, and its rate (HIP events over 20 copies after 3
    warm-up copies; bytes read + written) is the achievable HBM bandwidth of a streaming kernel."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    t0.record()
    for _ in range(reps):
        b.copy_(a)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / reps
    gbs = 2.0 * n * 4 / (ms * 1e-3) / 1e9
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3 + reps,
                      "ms_per_copy": ms, "gb_per_s": gbs, "frac_of_8tbs": gbs / HBM_PEAK_GBS}))


def vortex_ic_exact(xy, t):
    return W.vortex(xy, t=t)


if __name__ == "__main__":
    main()
