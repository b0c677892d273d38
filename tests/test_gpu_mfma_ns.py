"""The matrix-core Navier-Stokes stage (FP32, p = 3 and 4: face-trace launch + stage launch)
against the one-element-per-lane Navier-Stokes kernel and the FP64 oracle, with the reference's
face-gradient read and the plain one, on periodic and bounded meshes whose element counts are
not multiples of the 16-element wave tile or the 64-element block.

Tolerances as for the Euler matrix-core kernel: each FP32 kernel is held against the FP64
oracle (5e-5 of the residual's largest entry on meshes of a few dozen elements, 1e-4 at 4050
elements), and the matrix-core kernel's error may not exceed twice the per-lane kernel's (plus
1e-6).  After three RK3 steps the two kernels' states agree to relative 1e-6.
"""
import os

import numpy as np
import pytest

import dgx
import refcpu as R

pytestmark = pytest.mark.gpu

FAR = [1.0, 1.0, 0.0, 1 / 0.4 + 0.5]
MIXED = {"Bottom": ("far_field", FAR), "Top": ("slip_wall", [0, 0, 0, 0]),
         "Left": ("no_slip_wall", [1.0, 0.0, 0.0, 1.0])}  # Right: no condition


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


def _setup(p, case, face_grad="reference", n=5, mu=0.05):
    if case == "periodic":
        verts, tris = R.split_quad_mesh(n, 2 * np.pi)
        ora = R.Oracle(verts, tris, p, basis="dubiner", physics="ns", mu=mu, prandtl=0.71, sigma0=5.0,
                       ns_face_grad=face_grad)
        kw = {}
        xy = ora.volume_points()
        u0 = ora.project(R.taylor_green_state(xy))
    else:
        verts, tris = R.split_quad_mesh(n, 1.0)
        ora = R.Oracle(verts, tris, p, basis="dubiner", physics="ns", mu=mu, prandtl=0.71, sigma0=5.0,
                       ns_face_grad=face_grad, periodic=0)
        bcs = MIXED if case == "mixed" else {t: (case, FAR if case == "far_field" else [1.0, 0.0, 0.0, 1.0])
                                             for t in ("Bottom", "Right", "Top", "Left")}
        for tag, (typ, val) in bcs.items():
            ora.set_bc(tag, typ, val)
        kw = dict(periodic=0, bcs=bcs)
        xy = ora.volume_points()
        u0 = ora.project(R.vortex_state(xy, xc=0.5, yc=0.5, beta=1.0, length=1.0))
    # element-wise perturbation: every gradient and face jump is exercised
    rng = np.random.default_rng(p)
    u0 = u0 + 1e-3 * rng.standard_normal(u0.shape) * np.abs(u0).max(axis=(0, 1))
    wf = dgx.NavierStokesWeakFormulation(1.4, mu, 0.71, 5.0, face_grad=face_grad)
    mf = _context(False, verts, tris, p, weak_form=wf, basis="dubiner", precision="fp32", **kw)
    ln = _context(True, verts, tris, p, weak_form=wf, basis="dubiner", precision="fp32", **kw)
    return ora, mf, ln, u0


def test_ns_matrix_core_kernel_is_selected():
    ora, mf, ln, u0 = _setup(3, "periodic")
    assert mf.info.stage_kernel == 1 and ln.info.stage_kernel == 0
    verts, tris = R.split_quad_mesh(3, 1.0)
    wf = dgx.NavierStokesWeakFormulation(1.4, 0.01, 0.71, 5.0)
    # FP64 and orders without matrix-core tables keep the per-lane kernel
    assert dgx.Context(verts, tris, 3, weak_form=wf, precision="fp64").info.stage_kernel == 0
    assert dgx.Context(verts, tris, 2, weak_form=wf, precision="fp32").info.stage_kernel == 0


@pytest.mark.parametrize("p", [3, 4])
@pytest.mark.parametrize("face_grad", ["reference", "correct"])
@pytest.mark.parametrize("case", ["periodic", "far_field", "no_slip_wall", "mixed"])
def test_ns_mfma_residual_matches_lane_kernel_and_oracle(p, face_grad, case):
    ora, mf, ln, u0 = _setup(p, case, face_grad)
    assert mf.K % 16 != 0
    r_mf, r_ln, r_ref = mf.residual(u0), ln.residual(u0), ora.residual(u0)
    e_mf, e_ln = _rel(r_mf, r_ref), _rel(r_ln, r_ref)
    assert e_mf < 5e-5 and e_ln < 5e-5, (e_mf, e_ln)
    assert e_mf < 2 * e_ln + 1e-6, (e_mf, e_ln)


@pytest.mark.parametrize("p", [3, 4])
@pytest.mark.parametrize("case", ["periodic", "mixed"])
def test_ns_mfma_residual_with_storage_offset(p, case):
    # Once a state is set the FP32 state carries a storage offset, and the matrix-core kernel
    # forms its fluxes from the deviation, with the momentum pressure relative to a constant;
    # every face must carry that constant's flux, the mixed box's face without a condition too.
    ora, mf, ln, u0 = _setup(p, case)
    for c in (mf, ln):
        c.set_state(u0)
    r_mf, r_ln, r_ref = mf.residual(u0), ln.residual(u0), ora.residual(u0)
    e_mf, e_ln = _rel(r_mf, r_ref), _rel(r_ln, r_ref)
    assert e_mf < 5e-5 and e_ln < 5e-5, (e_mf, e_ln)
    assert e_mf < 2 * e_ln + 1e-6, (e_mf, e_ln)


@pytest.mark.parametrize("p", [3, 4])
def test_ns_mfma_viscous_part_matches_oracle(p):
    # the viscous contribution alone (NS residual minus the Euler residual of the same state)
    # is held to the same bar, so a wrong viscous term cannot hide under the inviscid one
    ora, mf, ln, u0 = _setup(p, "mixed", "reference", mu=0.05)
    verts, tris = R.split_quad_mesh(5, 1.0)
    eo = R.Oracle(verts, tris, p, basis="dubiner", periodic=0)
    for tag, (typ, val) in MIXED.items():
        eo.set_bc(tag, typ, val)
    ec = _context(False, verts, tris, p, basis="dubiner", precision="fp32", periodic=0, bcs=MIXED)
    v_ref = ora.residual(u0) - eo.residual(u0)
    v_mf = mf.residual(u0) - ec.residual(u0)
    assert np.abs(v_ref).max() > 1e-3 * np.abs(eo.residual(u0)).max()
    assert np.abs(v_mf - v_ref).max() < 5e-5 * np.abs(ora.residual(u0)).max()


# (the mixed box's unconditioned side and zero-valued slip-wall penalty state blow up within
# three steps at p = 4 in the oracle as well: the steps run on the far-field box)
@pytest.mark.parametrize("p", [3, 4])
@pytest.mark.parametrize("case", ["periodic", "far_field"])
def test_ns_mfma_rk3_steps_match_lane_kernel(p, case):
    ora, mf, ln, u0 = _setup(p, case, n=7, mu=0.01)
    ln.set_state(u0)
    dt = 0.2 * dgx.stable_dt(ln, 0.3)
    assert dt > 0
    for c in (mf, ln):
        c.set_state(u0)
        c.rk3_steps(dt, 3)
    a, b = mf.get_state(), ln.get_state()
    assert _rel(a, b) < 1e-6
    u_ref, _ = ora.rk3(u0, dt, 3)
    assert _rel(a, u_ref) < 5e-5


def test_ns_mfma_large_mesh_tail_block():
    # 2 * 45^2 = 4050 elements: 63 full blocks of 64 and a partial one
    verts, tris = R.split_quad_mesh(45, 2 * np.pi)
    ora = R.Oracle(verts, tris, 3, basis="dubiner", physics="ns", mu=0.01, prandtl=0.71, sigma0=5.0,
                   ns_face_grad="reference")
    u0 = ora.project(R.taylor_green_state(ora.volume_points()))
    # the smooth state's residual is a small difference of large terms: perturb it per element
    rng = np.random.default_rng(7)
    u0 = u0 + 1e-3 * rng.standard_normal(u0.shape) * np.abs(u0).max(axis=(0, 1))
    wf = dgx.NavierStokesWeakFormulation(1.4, 0.01, 0.71, 5.0, face_grad="reference")
    mf = _context(False, verts, tris, 3, weak_form=wf, basis="dubiner", precision="fp32")
    ln = _context(True, verts, tris, 3, weak_form=wf, basis="dubiner", precision="fp32")
    r_ref = ora.residual(u0)
    r_mf = mf.residual(u0)
    e_mf, e_ln = _rel(r_mf, r_ref), _rel(ln.residual(u0), r_ref)
    assert e_mf < 1e-4 and e_ln < 1e-4, (e_mf, e_ln)
    assert e_mf < 2 * e_ln + 1e-6, (e_mf, e_ln)
    tail = slice(mf.K - 50, mf.K)
    assert np.abs(r_mf[tail] - r_ref[tail]).max() < 1e-4 * np.abs(r_ref).max()
