"""Boundary conditions given as functions of position (dgx_set_bc_function; the reference's
BoundaryConditionEuler with a function, S/boundary/conditions.cpp:47-60) on every stage kernel,
against the FP64 oracle with the same functions: far-field states (conserved variables) and a
moving no-slip wall (primitive variables) that vary along the boundary, with the Euler boundary
flux evaluated at the reference's row-major read of the face points and at the points themselves.

Tolerances: FP64 per-lane kernels 1e-11 of the residual's largest entry (rounding only); FP32
kernels 5e-5, as for constant conditions (the residual is a small difference of larger terms).
"""
import os

import numpy as np
import pytest

import dgx
import refcpu as R

pytestmark = pytest.mark.gpu
GAM = 1.4
TAGS = ("Bottom", "Right", "Top", "Left")


def _cons(rho, u, v, p):
    return [rho, rho * u, rho * v, p / (GAM - 1) + 0.5 * rho * (u * u + v * v)]


def far_fn(x, y):
    return _cons(1.0 + 0.2 * x, 0.8 + 0.3 * y, 0.1 * x - 0.05, 1.0 + 0.4 * y)


def lid_fn(x, y):
    # moving wall: rho, u = 4 x (1 - x) (the lid profile), v = 0, p
    return [1.0 + 0.1 * x, 4.0 * x * (1.0 - x), 0.0, 1.0]


BCS = {"Bottom": ("far_field", far_fn), "Top": ("no_slip_wall", lid_fn),
       "Left": ("far_field", far_fn), "Right": ("slip_wall", [0, 0, 0, 0])}


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


def _wf(physics):
    return (dgx.NavierStokesWeakFormulation(GAM, 0.02, 0.71, 5.0, face_grad="reference")
            if physics == "ns" else dgx.EulerWeakFormulation(GAM))


def _oracle(verts, tris, p, physics, points, bcs=BCS):
    kw = dict(physics="ns", mu=0.02, prandtl=0.71, sigma0=5.0, ns_face_grad="reference") if physics == "ns" else {}
    ora = R.Oracle(verts, tris, p, basis="dubiner", periodic=0, euler_bc_points=points, **kw)
    for tag, (typ, val) in bcs.items():
        if callable(val):
            ora.set_bc(tag, typ, fn=val)
        else:
            ora.set_bc(tag, typ, val)
    return ora


def _state(ora, seed):
    xy = ora.volume_points()
    u0 = ora.project(R.vortex_state(xy, xc=0.45, yc=0.55, beta=1.0, length=1.0))
    rng = np.random.default_rng(seed)
    return u0 + 1e-3 * rng.standard_normal(u0.shape) * np.abs(u0).max(axis=(0, 1))


@pytest.mark.parametrize("physics", ["euler", "ns"])
@pytest.mark.parametrize("points", ["reference", "correct"])
@pytest.mark.parametrize("p", [2, 3, 4])
def test_function_bcs_match_oracle_on_every_kernel(physics, points, p):
    verts, tris = R.split_quad_mesh(5, 1.0)
    ora = _oracle(verts, tris, p, physics, points)
    u0 = _state(ora, p)
    r_ref = ora.residual(u0)
    kw = dict(weak_form=_wf(physics), basis="dubiner", periodic=0, bcs=BCS, euler_bc_points=points)
    c64 = _context(False, verts, tris, p, precision="fp64", **kw)
    assert _rel(c64.residual(u0), r_ref) < 1e-11
    mf = _context(False, verts, tris, p, precision="fp32", **kw)
    ln = _context(True, verts, tris, p, precision="fp32", **kw)
    # matrix cores: Euler at every order, Navier-Stokes at p = 3, 4
    assert mf.info.stage_kernel == (1 if physics == "euler" or p >= 3 else 0) and ln.info.stage_kernel == 0
    e_mf, e_ln = _rel(mf.residual(u0), r_ref), _rel(ln.residual(u0), r_ref)
    assert e_mf < 5e-5 and e_ln < 5e-5, (e_mf, e_ln)


@pytest.mark.parametrize("physics", ["euler", "ns"])
def test_the_two_point_reads_differ_as_the_oracle_does(physics):
    # the reference's read evaluates the condition away from the face: the residuals differ
    # by more than rounding, and by what the oracle's two reads differ by
    verts, tris = R.split_quad_mesh(4, 1.0)
    u0 = None
    got, want = [], []
    for points in ("reference", "correct"):
        ora = _oracle(verts, tris, 3, physics, points)
        u0 = _state(ora, 9) if u0 is None else u0
        ctx = dgx.Context(verts, tris, 3, weak_form=_wf(physics), basis="dubiner", precision="fp64",
                          periodic=0, bcs=BCS, euler_bc_points=points)
        got.append(ctx.residual(u0))
        want.append(ora.residual(u0))
    d_want = want[0] - want[1]
    assert np.abs(d_want).max() > 1e-3 * np.abs(want[1]).max()
    assert np.abs((got[0] - got[1]) - d_want).max() < 1e-11 * np.abs(want[1]).max()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_constant_function_equals_constant_value(precision):
    verts, tris = R.split_quad_mesh(4, 1.0)
    far = far_fn(0.3, 0.7)
    wall = lid_fn(0.5, 0.0)
    const = {"Bottom": ("far_field", far), "Top": ("no_slip_wall", wall), "Left": ("far_field", far)}
    fns = {"Bottom": ("far_field", lambda x, y: far), "Top": ("no_slip_wall", lambda x, y: wall),
           "Left": ("far_field", lambda x, y: far)}
    ora = _oracle(verts, tris, 3, "ns", "reference", const)
    u0 = _state(ora, 4)
    res = []
    for bcs in (const, fns):
        ctx = dgx.Context(verts, tris, 3, weak_form=_wf("ns"), basis="dubiner", precision=precision,
                          periodic=0, bcs=bcs)
        res.append(ctx.residual(u0))
    assert np.array_equal(res[0], res[1])


def test_function_bcs_rk3_steps_on_matrix_cores():
    verts, tris = R.split_quad_mesh(6, 1.0)
    ora = _oracle(verts, tris, 3, "euler", "correct")
    u0 = _state(ora, 5)
    ctx = dgx.Context(verts, tris, 3, basis="dubiner", precision="fp32", periodic=0, bcs=BCS,
                      euler_bc_points="correct")
    assert ctx.info.stage_kernel == 1
    ctx.set_state(u0)
    dt = 0.5 * dgx.stable_dt(ctx, 0.3)
    ctx.rk3_steps(dt, 3)
    u_ref, _ = ora.rk3(u0, dt, 3)
    assert _rel(ctx.get_state(), u_ref) < 2e-5


def test_function_bcs_adjoint_against_oracle_fd():
    # FP64 vector-Jacobian product of the residual with position-dependent far-field and
    # no-slip ghosts, by central differences of the oracle residual
    verts, tris = R.split_quad_mesh(4, 1.0)
    bcs = {"Bottom": ("far_field", far_fn), "Left": ("no_slip_wall", lid_fn)}
    ora = _oracle(verts, tris, 2, "euler", "reference", bcs)
    ctx = dgx.Context(verts, tris, 2, basis="dubiner", precision="fp64", periodic=0, bcs=bcs)
    u0 = _state(ora, 11)
    lam = np.random.default_rng(12).standard_normal(u0.shape)
    g = ctx.residual_vjp(lam, u0)
    ok = 0
    for d in [np.random.default_rng(30 + k).standard_normal(u0.shape) for k in range(4)]:
        errs = []
        for h in (1e-6, 1e-7):
            fd = np.sum((ora.residual(u0 + h * d) - ora.residual(u0 - h * d)) / (2 * h) * lam)
            errs.append(abs(np.sum(g * d) - fd) / abs(fd))
        ok += min(errs) < 1e-6
    assert ok >= 3


def test_function_bcs_matrix_core_adjoint_matches_lane():
    verts, tris = R.split_quad_mesh(6, 1.0)
    # the adjoint differentiates far-field and no-slip ghosts (slip walls are rejected)
    bcs = {t: v for t, v in BCS.items() if v[0] != "slip_wall"}
    ora = _oracle(verts, tris, 3, "euler", "reference", bcs)
    u0 = _state(ora, 13)
    lam = np.random.default_rng(14).standard_normal(u0.shape)
    kw = dict(basis="dubiner", precision="fp32", periodic=0, bcs=bcs)
    g = [_context(lane, verts, tris, 3, **kw).residual_vjp(lam, u0) for lane in (False, True)]
    assert np.linalg.norm(g[0] - g[1]) / np.linalg.norm(g[1]) < 1e-5


def test_set_bc_function_rejects_bad_index():
    verts, tris = R.split_quad_mesh(3, 1.0)
    ctx = dgx.Context(verts, tris, 2, precision="fp64", periodic=0, bcs=BCS)
    cb = dgx.BC_FN(lambda x, y, out, user: None)
    for k in (-1, 4):
        with pytest.raises(dgx.DGXError):
            dgx._check(dgx.load().dgx_set_bc_function(ctx._h, k, cb, None))
    with pytest.raises(ValueError):
        dgx.Context(verts, tris, 2, periodic=0, bcs=BCS, euler_bc_points="elsewhere")
