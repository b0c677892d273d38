"""The matrix-core stage kernel (FP32 Euler, p = 1 to 4) against the one-element-per-lane
stage kernel and the FP64 oracle, on periodic and bounded meshes whose element counts are not
multiples of the 16-element wave tile or the 64-element block.

Tolerances: the residual is a small difference of volume and face terms that are each larger
by about the inverse mesh width, so FP32 rounding shows up at 1e-6 to 1e-5 of its largest entry
and differently in the two kernels' summation orders.  Each kernel is therefore held against the
FP64 oracle (2e-5 on meshes of a few dozen elements, 1e-4 at 4050 elements) and the
matrix-core kernel's error may not exceed twice the per-lane kernel's (plus 1e-6).  After three
RK3 steps the two kernels' states agree to relative 1e-6.
"""
import os

import numpy as np
import pytest

import dgx
import refcpu as R

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _context(lane, *args, **kw):
    # DGX_STAGE_KERNEL is read when the context is created
    if lane:
        os.environ["DGX_STAGE_KERNEL"] = "lane"
    try:
        return dgx.Context(*args, **kw)
    finally:
        os.environ.pop("DGX_STAGE_KERNEL", None)


FAR = [1.0, 1.0, 0.0, 1 / 0.4 + 0.5]
MIXED = {"Bottom": ("far_field", FAR), "Top": ("slip_wall", [0, 0, 0, 0]),
         "Left": ("no_slip_wall", [1.0, 0.0, 0.0, 1.0])}  # Right: no condition


def _setup(p, case, n=5):
    length = 10.0 if case == "periodic" else 1.0
    verts, tris = R.split_quad_mesh(n, length)
    if case == "periodic":
        ora = R.Oracle(verts, tris, p, basis="dubiner")
        kw = {}
        xy = ora.volume_points()
        pert = np.repeat(0.01 * R.uniform_pm1(3, np.arange(ora.K))[:, None], ora.Nq, axis=1)
        U = R.vortex_state(xy, rho_pert=pert)
    else:
        ora = R.Oracle(verts, tris, p, basis="dubiner", periodic=0)
        bcs = MIXED if case == "mixed" else {t: (case, FAR if case == "far_field" else
                                                 [1.0, 0.0, 0.0, 1.0] if case == "no_slip_wall" else [0, 0, 0, 0])
                                             for t in ("Bottom", "Right", "Top", "Left")}
        for tag, (typ, val) in bcs.items():
            ora.set_bc(tag, typ, val)
        kw = dict(periodic=0, bcs=bcs)
        xy = ora.volume_points()
        U = R.vortex_state(xy, xc=0.5, yc=0.5, beta=1.0, length=1.0)
    u0 = ora.project(U)
    mf = _context(False, verts, tris, p, basis="dubiner", precision="fp32", **kw)
    ln = _context(True, verts, tris, p, basis="dubiner", precision="fp32", **kw)
    return ora, mf, ln, u0


@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("case", ["periodic", "far_field", "slip_wall", "no_slip_wall", "mixed"])
def test_mfma_residual_matches_lane_kernel_and_oracle(p, case):
    ora, mf, ln, u0 = _setup(p, case)
    assert mf.K % 16 != 0
    r_mf, r_ln, r_ref = mf.residual(u0), ln.residual(u0), ora.residual(u0)
    e_mf, e_ln = _rel(r_mf, r_ref), _rel(r_ln, r_ref)
    assert e_mf < 2e-5 and e_ln < 2e-5, (e_mf, e_ln)
    assert e_mf < 2 * e_ln + 1e-6, (e_mf, e_ln)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("case", ["periodic", "mixed"])
def test_mfma_rk3_steps_match_lane_kernel(p, case):
    ora, mf, ln, u0 = _setup(p, case, n=7)
    ln.set_state(u0)
    dt = 0.2 * dgx.stable_dt(ln, 0.3)
    assert dt > 0
    for c in (mf, ln):
        c.set_state(u0)
        c.rk3_steps(dt, 3)
    a, b = mf.get_state(), ln.get_state()
    assert _rel(a, b) < 1e-6
    u_ref, _ = ora.rk3(u0, dt, 3)
    assert _rel(a, u_ref) < 2e-5


def test_mfma_large_mesh_tail_block():
    # 2 * 45^2 = 4050 elements: 63 full blocks of 64 and a partial one
    verts, tris = R.split_quad_mesh(45)
    ora = R.Oracle(verts, tris, 3, basis="dubiner")
    xy = ora.volume_points()
    u0 = ora.project(R.vortex_state(xy))
    mf = _context(False, verts, tris, 3, basis="dubiner", precision="fp32")
    ln = _context(True, verts, tris, 3, basis="dubiner", precision="fp32")
    r_ref = ora.residual(u0)
    e_mf, e_ln = _rel(mf.residual(u0), r_ref), _rel(ln.residual(u0), r_ref)
    assert e_mf < 1e-4 and e_ln < 1e-4, (e_mf, e_ln)
    assert e_mf < 2 * e_ln + 1e-6, (e_mf, e_ln)
    # the partial last block (its residual is free-stream small) against the global scale
    tail = slice(mf.K - 50, mf.K)
    assert np.abs(mf.residual(u0)[tail] - r_ref[tail]).max() < 1e-4 * np.abs(r_ref).max()


def test_p1_kernel_choice_by_mesh_size():
    # at p = 1 a single FP32 context keeps the per-lane kernel above 98304 elements (it is the
    # faster one there); each choice is held against the FP64 oracle like the per-lane kernel
    got = {}
    for n in (221, 222):  # K = 97682 (tiles), 98568 (per lane)
        verts, tris = R.split_quad_mesh(n, 10.0)
        ora = R.Oracle(verts, tris, 1, basis="dubiner")
        u0 = ora.project(R.vortex_state(ora.volume_points()))
        ctx = dgx.Context(verts, tris, 1, basis="dubiner", precision="fp32")
        ln = _context(True, verts, tris, 1, basis="dubiner", precision="fp32")
        got[n] = ctx.info.stage_kernel
        r_ref = ora.residual(u0)
        e, e_ln = _rel(ctx.residual(u0), r_ref), _rel(ln.residual(u0), r_ref)
        assert e < 1e-3 and e < 2 * e_ln + 1e-6, (n, e, e_ln)
        ctx.close()
        ln.close()
    assert got == {221: 1, 222: 0}


# ---------------- stage adjoint ------------------------------------------------------------

def _adj_setup(p, case, n=5):
    ora, mf, ln, u0 = _setup(p, "periodic" if case == "periodic" else case, n)
    verts, tris = R.split_quad_mesh(n, 10.0 if case == "periodic" else 1.0)
    kw = {} if case == "periodic" else dict(periodic=0, bcs=MIXED if case == "mixed" else {
        t: (case, FAR if case == "far_field" else [1.0, 0.0, 0.0, 1.0]) for t in ("Bottom", "Right", "Top", "Left")})
    c64 = dgx.Context(verts, tris, p, basis="dubiner", precision="fp64", **kw)
    return ora, mf, ln, c64, u0


ADJ_MIXED = {"Bottom": ("far_field", FAR), "Left": ("no_slip_wall", [1.0, 0.0, 0.0, 1.0])}


@pytest.mark.parametrize("p", [3, 4])
@pytest.mark.parametrize("case", ["periodic", "far_field", "no_slip_wall"])
def test_mfma_vjp_matches_lane_and_fp64(p, case):
    # the matrix-core adjoint against the per-lane FP32 adjoint and the FP64 adjoint, for the
    # residual (unfolded tables) and the right-hand side (folded tables)
    ora, mf, ln, c64, u0 = _adj_setup(p, case)
    lam = np.random.default_rng(p).standard_normal(u0.shape)
    for f in ("residual_vjp", "rhs_vjp"):
        g64 = getattr(c64, f)(lam, u0)
        e_mf = _rel(getattr(mf, f)(lam, u0), g64)
        e_ln = _rel(getattr(ln, f)(lam, u0), g64)
        assert e_mf < 5e-5 and e_ln < 5e-5, (f, e_mf, e_ln)
        assert e_mf < 2 * e_ln + 1e-6, (f, e_mf, e_ln)


@pytest.mark.parametrize("p", [3, 4])
def test_mfma_rk3_vjp_matches_lane(p):
    ora, mf, ln, u0 = _setup(p, "periodic", n=7)
    ln.set_state(u0)
    dt = dgx.stable_dt(ln, 0.3)
    lam_T = np.random.default_rng(3).standard_normal(u0.shape)
    g = []
    for c in (mf, ln):
        c.set_state(u0)
        g.append(c.rk3_vjp(dt, 6, 3, lam_T))
    assert np.linalg.norm(g[0] - g[1]) / np.linalg.norm(g[1]) < 1e-5


@pytest.mark.parametrize("n", [128, 132])
def test_mfma_vjp_both_launch_forms(n):
    # A p = 3 stage adjoint grid of at most two tiles per CU (n = 128: 512 tiles, two per CU of
    # an MI355X) launches the kernel's latency form, a larger one (n = 132: 545 tiles) the other
    # form.  Both against the FP64 adjoint, beside the per-lane FP32 adjoint: the stage adjoints
    # (residual and right-hand side) to 1e-6 of their largest entry (measured 2-3e-7), and the
    # trajectory adjoint over two RK3 steps with a checkpoint per step, whose FP32 error is that
    # of the FP32 trajectory (6e-5 relative L2 for both kernels, most of it in rho v) and is held
    # to 2e-4 and to twice the per-lane kernel's.
    verts, tris = R.split_quad_mesh(n, 10.0)
    mf = _context(False, verts, tris, 3, basis="dubiner", precision="fp32")
    ln = _context(True, verts, tris, 3, basis="dubiner", precision="fp32")
    c64 = _context(False, verts, tris, 3, basis="dubiner", precision="fp64")
    assert mf.info.stage_kernel == 1 and ln.info.stage_kernel == 0
    pert = np.repeat(0.01 * R.uniform_pm1(5, np.arange(mf.K))[:, None], mf.Nq, axis=1)
    c64.project(R.vortex_state(c64.volume_points(), rho_pert=pert))
    u0 = c64.get_state()
    lam = np.random.default_rng(3).standard_normal(u0.shape)
    for f in ("residual_vjp", "rhs_vjp"):
        g64 = getattr(c64, f)(lam, u0)
        e_mf = _rel(getattr(mf, f)(lam, u0), g64)
        e_ln = _rel(getattr(ln, f)(lam, u0), g64)
        assert e_mf < 1e-6 and e_mf < 2 * e_ln + 1e-7, (f, e_mf, e_ln)
    dt = dgx.stable_dt(c64, 0.3)
    lam_T = np.random.default_rng(9).standard_normal(u0.shape)
    g = []
    for c in (mf, ln, c64):
        c.set_state(u0)
        g.append(c.rk3_vjp(dt, 2, 1, lam_T))
    e_mf, e_ln = (np.linalg.norm(x - g[2]) / np.linalg.norm(g[2]) for x in g[:2])
    assert e_mf < 2e-4 and e_mf < 2 * e_ln, (e_mf, e_ln)
    for c in (mf, ln, c64):
        c.close()


def test_vjp_no_slip_ghost_against_oracle_fd():
    # FP64 adjoint on a bounded mesh with far-field and no-slip faces, by central differences
    # of the oracle residual: the no-slip ghost state is built from primitive values
    verts, tris = R.split_quad_mesh(4, 1.0)
    ora = R.Oracle(verts, tris, 2, basis="dubiner", periodic=0)
    for tag, (typ, val) in ADJ_MIXED.items():
        ora.set_bc(tag, typ, val)
    ctx = dgx.Context(verts, tris, 2, basis="dubiner", precision="fp64", periodic=0, bcs=ADJ_MIXED)
    xy = ora.volume_points()
    U = R.vortex_state(xy, xc=0.43, yc=0.57, beta=1.0, length=1.0)
    rho, u, v = U[..., 0], U[..., 1] / U[..., 0], U[..., 2] / U[..., 0] + 0.3
    pr = 0.4 * (U[..., 3] - 0.5 * (U[..., 1] ** 2 + U[..., 2] ** 2) / rho)
    u0 = ora.project(np.stack([rho, rho * u, rho * v, pr / 0.4 + 0.5 * rho * (u * u + v * v)], axis=-1))
    lam = np.random.default_rng(11).standard_normal(u0.shape)
    g = ctx.residual_vjp(lam, u0)
    ok = 0
    for d in [np.random.default_rng(20 + k).standard_normal(u0.shape) for k in range(4)]:
        errs = []
        for h in (1e-6, 1e-7):
            fd = np.sum((ora.residual(u0 + h * d) - ora.residual(u0 - h * d)) / (2 * h) * lam)
            errs.append(abs(np.sum(g * d) - fd) / abs(fd))
        ok += min(errs) < 1e-6
    assert ok >= 3


# ---------------- bfloat16 volume products ---------------------------------------------------

@pytest.mark.parametrize("p", [3, 4])
def test_bf16_volume_variant_error_is_bf16_sized(p):
    # bfloat16 keeps 8 significant bits (relative rounding 2^-9) of the volume operands, so the
    # variant's residual error scales with the volume term (about 100x the residual here, where
    # volume and face terms nearly cancel): above the FP32 kernel's, below 2^-5 of that term
    ora, mf, ln, u0 = _setup(p, "periodic", n=6)
    verts, tris = R.split_quad_mesh(6)
    bf = dgx.Context(verts, tris, p, basis="dubiner", precision="fp32_bf16vol")
    assert bf.info.stage_kernel == 1
    r_ref = ora.residual(u0)
    vol = max(np.abs(ora.volume_residual(e, u0[e])).max() for e in range(ora.K))
    d_bf = np.abs(bf.residual(u0) - r_ref).max()
    d_32 = np.abs(mf.residual(u0) - r_ref).max()
    assert d_32 < d_bf < 2 ** -5 * vol, (d_bf, d_32, vol)


def test_bf16_volume_variant_steps_and_conserves():
    ora, mf, ln, u0 = _setup(3, "periodic", n=8)
    verts, tris = R.split_quad_mesh(8)
    bf = dgx.Context(verts, tris, 3, basis="dubiner", precision="fp32_bf16vol")
    bf.set_state(u0)
    mf.set_state(u0)
    s0 = bf.conservation_sums()
    dt = dgx.stable_dt(mf, 0.3)
    bf.rk3_steps(dt, 20)
    mf.rk3_steps(dt, 20)
    a, b = bf.get_state(), mf.get_state()
    assert np.isfinite(a).all()
    assert np.linalg.norm(a[..., 0] - b[..., 0]) / np.linalg.norm(b[..., 0]) < 1e-2
    # the fluxes stay a telescoping sum over faces: mass is conserved to FP32 rounding
    assert abs(bf.conservation_sums()[0] - s0[0]) < 1e-5 * abs(s0[0])


def test_bf16_volume_variant_limits():
    verts, tris = R.split_quad_mesh(4)
    for kw in (dict(order=2), dict(order=3, weak_form=dgx.NavierStokesWeakFormulation())):
        order = kw.pop("order")
        with pytest.raises(dgx.DGXError) as e:
            dgx.Context(verts, tris, order, precision="fp32_bf16vol", **kw)
        assert e.value.status == dgx.DGX_E_UNSUPPORTED
    bf = dgx.Context(verts, tris, 3, precision="fp32_bf16vol")
    with pytest.raises(dgx.DGXError) as e:
        bf.rhs_vjp(np.zeros((bf.K, bf.Np, 4)), np.zeros((bf.K, bf.Np, 4)) + 1.0)
    assert e.value.status == dgx.DGX_E_UNSUPPORTED
