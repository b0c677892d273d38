"""Quadrilateral elements on the GPU (Legendre tensor basis, tensor Gauss-Legendre rule,
bilinear mapping, per-element inverse mass) against the FP64 oracle's quad space:

* mesh helpers: the structured quad generator and the quad connectivity equal the oracle's;
* residuals on distorted periodic and bounded quads (Euler, Navier-Stokes with both
  face-gradient reads and both Legendre endpoint derivatives, constant and position-dependent
  conditions), p = 1..4: FP64 within 1e-11 of the residual's largest entry, FP32 within 5e-5;
* RK3 trajectories, projection, evaluation, conservation sums, CFL, density range, L2 error;
* the reference's shear-flow and compressible-Blasius accuracy runs on quads (FP64, error
  norms equal to the oracle's, and that test's bounds with the true boundary points).
"""
import numpy as np
import pytest

import dgx
import flows
import refcpu as R

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _distorted(n, length, seed, amp=0.15):
    verts, quads = R.rect_quad_mesh(0.0, length, 0.0, length, length / n)
    rng = np.random.default_rng(seed)
    inner = (verts[:, 0] > 0) & (verts[:, 0] < length) & (verts[:, 1] > 0) & (verts[:, 1] < length)
    verts = verts.copy()
    verts[inner] += rng.uniform(-amp, amp, (inner.sum(), 2)) * (length / n)
    return verts, quads


def test_mesh_helpers_match_the_oracle():
    for args in [(0.0, 1.0, 0.0, 1.0, 0.05), (0.05, 0.55, 0.0, 0.2, 0.02), (-1.0, 2.0, 0.5, 1.25, 0.25)]:
        v, q = dgx.rect_quad_mesh(*args)
        rv, rq = R.rect_quad_mesh(*args)
        assert np.array_equal(q, rq) and np.abs(v - rv).max() < 1e-15
    verts, quads = _distorted(5, 2.0, 1)
    for periodic in (0, 1, 2, 3):
        nb = dgx.connect(verts, quads, periodic)
        ora = R.Oracle(verts, quads, 1, periodic=periodic)
        on = ora.neighbors()
        inner = on[..., 0] >= 0
        assert np.array_equal(nb[..., 0] >= 0, inner)
        assert np.array_equal(nb[inner], on[inner])


FAR = flows.prim2cons([1.0, 0.6, 0.2, 1.0])


def _far_fn(x, y):
    return flows.prim2cons([1.0 + 0.1 * x, 0.6 + 0.2 * y, 0.2 - 0.1 * x, 1.0 + 0.05 * y])


def _lid(x, y):
    return [1.0, 4.0 * x * (1.0 - x), 0.0, 1.0]


CASES = {
    "periodic": (3, {}),
    "far_field": (0, {t: ("far_field", FAR) for t in ("Bottom", "Right", "Top", "Left")}),
    "mixed_functions": (0, {"Bottom": ("far_field", _far_fn), "Top": ("no_slip_wall", _lid),
                            "Left": ("slip_wall", [0, 0, 0, 0])}),  # Right: no condition
}


def _pair(p, case, physics="euler", face_grad="reference", endpoint="reference", points="reference", n=4):
    periodic, bcs = CASES[case]
    verts, quads = _distorted(n, 1.0, p)
    kw = dict(physics="ns", mu=0.02, prandtl=0.71, sigma0=5.0, ns_face_grad=face_grad) if physics == "ns" else {}
    ora = R.Oracle(verts, quads, p, periodic=periodic, euler_bc_points=points, legendre_endpoint=endpoint, **kw)
    for tag, (typ, val) in bcs.items():
        if callable(val):
            ora.set_bc(tag, typ, fn=val)
        else:
            ora.set_bc(tag, typ, val)
    wf = (dgx.NavierStokesWeakFormulation(1.4, 0.02, 0.71, 5.0, face_grad=face_grad) if physics == "ns"
          else dgx.EulerWeakFormulation(1.4))
    ctx = {prec: dgx.Context(verts, quads, p, weak_form=wf, precision=prec, periodic=periodic, bcs=bcs,
                             euler_bc_points=points, legendre_endpoint=endpoint) for prec in ("fp64", "fp32")}
    xy = ora.volume_points()
    u0 = ora.project(R.vortex_state(xy, xc=0.45, yc=0.55, beta=1.0, length=1.0))
    rng = np.random.default_rng(10 + p)
    u0 = u0 + 1e-3 * rng.standard_normal(u0.shape) * np.abs(u0).max(axis=(0, 1))
    return ora, ctx, u0


@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("case", list(CASES))
def test_euler_residual_on_quads(p, case):
    ora, ctx, u0 = _pair(p, case)
    assert ctx["fp64"].info.n_modes == (p + 1) ** 2 and ctx["fp64"].info.n_volume_points == (p + 1) ** 2
    r_ref = ora.residual(u0)
    assert _rel(ctx["fp64"].residual(u0), r_ref) < 1e-11
    assert _rel(ctx["fp32"].residual(u0), r_ref) < 5e-5


@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("face_grad", ["reference", "correct"])
@pytest.mark.parametrize("endpoint", ["reference", "correct"])
def test_ns_residual_on_quads(p, face_grad, endpoint):
    for case in ("periodic", "mixed_functions"):
        ora, ctx, u0 = _pair(p, case, "ns", face_grad, endpoint)
        r_ref = ora.residual(u0)
        assert _rel(ctx["fp64"].residual(u0), r_ref) < 1e-11, case
        assert _rel(ctx["fp32"].residual(u0), r_ref) < 5e-5, case


def test_legendre_endpoint_changes_the_ns_face_terms_like_the_oracle():
    got, want = [], []
    for endpoint in ("reference", "correct"):
        ora, ctx, u0 = _pair(2, "periodic", "ns", "correct", endpoint)
        got.append(ctx["fp64"].residual(u0))
        want.append(ora.residual(u0))
    d = want[0] - want[1]
    assert np.abs(d).max() > 1e-6 * np.abs(want[1]).max()
    assert np.abs((got[0] - got[1]) - d).max() < 1e-11 * np.abs(want[1]).max()


# the mixed case with Navier-Stokes terms blows up within three steps in the oracle too (the
# unconditioned Right boundary): Navier-Stokes steps on the far-field box
@pytest.mark.parametrize("physics,case", [("euler", "mixed_functions"), ("ns", "far_field")])
def test_rk3_and_diagnostics_on_quads(physics, case):
    ora, ctx, u0 = _pair(2, case, physics, n=5)
    c64, c32 = ctx["fp64"], ctx["fp32"]
    c64.set_state(u0)
    dt = 0.5 * dgx.stable_dt(c64, 0.3)
    assert abs(c64.max_cfl(dt) - ora.max_cfl(u0, dt)) < 1e-12 * ora.max_cfl(u0, dt)
    lo, hi = c64.density_range()
    assert np.allclose([lo, hi], ora.density_range(u0), rtol=1e-13, atol=0)
    assert np.allclose(c64.conservation_sums(), ora.conservation_sums(u0), rtol=1e-12, atol=1e-14)
    u_ref, _ = ora.rk3(u0, dt, 5)
    for c, tol in ((c64, 1e-11), (c32, 2e-5)):
        c.set_state(u0)
        c.rk3_steps(dt, 5)
        assert _rel(c.get_state(), u_ref) < tol
    # the inverse mass
    r = ora.residual(u0)
    assert _rel(c64.apply_minv(r), ora.apply_minv(r)) < 1e-12


def test_projection_and_evaluation_on_quads():
    verts, quads = _distorted(4, 1.0, 7)
    ora = R.Oracle(verts, quads, 3, periodic=3)
    c = dgx.Context(verts, quads, 3, precision="fp64")
    xy = c.volume_points()
    assert np.abs(xy - ora.volume_points()).max() < 1e-15
    Uq = R.vortex_state(xy, xc=0.5, yc=0.5, beta=1.0, length=1.0)
    c.project(Uq)
    u_ref = ora.project(Uq)
    assert _rel(c.get_state(), u_ref) < 1e-12
    assert _rel(c.eval_volume(), ora.eval_volume(u_ref)) < 1e-13
    # with (p+1)^2 Gauss points and modes the projection interpolates the point values
    rho = Uq[..., 0]
    W = np.stack([rho, Uq[..., 1] / rho, Uq[..., 2] / rho,
                  0.4 * (Uq[..., 3] - 0.5 * (Uq[..., 1] ** 2 + Uq[..., 2] ** 2) / rho)], axis=-1)
    assert np.all(c.l2_error(W, primitive=True) < 1e-13)
    # relative L2 error (w_q |J_q| weights) against a different field
    W2 = W * (1.0 + 0.01 * np.sin(3.0 * xy[..., :1] + 2.0 * xy[..., 1:]))
    err = c.l2_error(W2, primitive=True)
    want = flows.l2_errors_at_points(ora, u_ref, W2)
    assert np.all(want > 1e-4) and np.all(np.abs(err - want) <= 1e-9 * want), (err, want)


@pytest.mark.parametrize("points", ["reference", "correct"])
def test_reference_shear_flow_run_on_quads(points):
    verts, quads = dgx.rect_quad_mesh(0.0, 1.0, 0.0, 1.0, 0.05)
    ora, u0 = flows.ns_oracle(verts, quads, 2, 5e-6, 1e4, flows.shear_primitive, points)
    bcs = {t: ("far_field", lambda x, y: flows.prim2cons(flows.shear_primitive(x, y)))
           for t in ("Bottom", "Right", "Top", "Left")}
    wf = dgx.NavierStokesWeakFormulation(1.4, 5e-6, 0.72, 1e4)
    c = dgx.Context(verts, quads, 2, weak_form=wf, precision="fp64", periodic=0, bcs=bcs, euler_bc_points=points)
    c.set_state(u0)
    c.rk3_steps(1e-5, 50)
    u_ref, _ = ora.rk3(u0, 1e-5, 50)
    assert _rel(c.get_state(), u_ref) < 1e-11
    W = np.array([[flows.shear_primitive(*q) for q in e] for e in c.volume_points()])
    got = c.l2_error(W, primitive=True)
    want = flows.l2_errors(ora, u_ref, flows.shear_primitive)
    assert np.all(np.abs(got - want) <= 1e-4 * want + 1e-12), (got, want)
    if points == "correct":
        assert got[0] < 5e-4 and got[1] < 5e-3 and got[2] < 5e-4 and got[3] < 5e-3, got


def test_reference_blasius_run_on_quads():
    verts, quads = dgx.rect_quad_mesh(0.05, 0.55, 0.0, 0.2, 0.02)
    ora, u0 = flows.ns_oracle(verts, quads, 2, flows.BLASIUS["mu"], 1e4, flows.blasius_primitive)
    bcs = {t: ("far_field", lambda x, y: flows.prim2cons(flows.blasius_primitive(x, y)))
           for t in ("Bottom", "Right", "Top", "Left")}
    wf = dgx.NavierStokesWeakFormulation(1.4, flows.BLASIUS["mu"], 0.72, 1e4)
    for prec, tol in (("fp64", 1e-11), ("fp32", 1e-5)):
        c = dgx.Context(verts, quads, 2, weak_form=wf, precision=prec, periodic=0, bcs=bcs)
        c.set_state(u0)
        c.rk3_steps(2e-6, 100)
        u_ref, _ = ora.rk3(u0, 2e-6, 100)
        assert _rel(c.get_state(), u_ref) < tol, prec
        W = np.array([[flows.blasius_primitive(*q) for q in e] for e in c.volume_points()])
        got = c.l2_error(W, primitive=True)
        assert got[0] < 1e-3 and got[1] < 5e-3 and got[2] < 8e-3 and got[3] < 5e-3, (prec, got)
        if prec == "fp64":
            want = flows.l2_errors(ora, u_ref, flows.blasius_primitive)
            assert np.all(np.abs(got - want) <= 1e-6 * want), (got, want)


def test_quads_outside_the_built_operators_are_rejected(tmp_path):
    verts, quads = dgx.rect_quad_mesh(0.0, 1.0, 0.0, 1.0, 0.25)
    c = dgx.Context(verts, quads, 2, precision="fp64")
    u = np.zeros((c.K, c.Np, 4)) + 1.0
    for call in (lambda: c.residual_vjp(u, u), lambda: c.geometry()):
        with pytest.raises(dgx.DGXError) as e:
            call()
        assert e.value.status == dgx.DGX_E_UNSUPPORTED
    with pytest.raises(dgx.DGXError):
        dgx.Context(verts, quads, 2, basis="dubiner", precision="fp64")
    tv, tt = dgx.split_quad_mesh(2, 1.0)
    with pytest.raises(dgx.DGXError):
        dgx.Context(tv, tt, 2, basis="legendre", precision="fp64")


def _quad_mesh(kind, n, seed):
    """rectangles: every element a parallelogram (diagonal mass); distorted: none; partial:
    every third interior vertex moved, so parallelograms and general quads share waves."""
    if kind == "distorted":
        return _distorted(n, 1.0, seed)
    verts, quads = R.rect_quad_mesh(0.0, 1.0, 0.0, 1.0, 1.0 / n)
    if kind == "partial":
        verts = verts.copy()
        inner = np.nonzero((verts[:, 0] > 0) & (verts[:, 0] < 1) & (verts[:, 1] > 0) & (verts[:, 1] < 1))[0][::3]
        verts[inner] += np.random.default_rng(seed).uniform(-0.15, 0.15, (len(inner), 2)) / n
    return verts, quads


def _lane_context(*args, **kw):
    import os
    os.environ["DGX_STAGE_KERNEL"] = "lane"  # read when the context is created
    try:
        return dgx.Context(*args, **kw)
    finally:
        os.environ.pop("DGX_STAGE_KERNEL", None)


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("kind", ["rectangles", "distorted", "partial"])
@pytest.mark.parametrize("case", ["periodic", "mixed_functions"])
def test_four_lane_kernel_against_the_oracle_and_the_lane_kernel(p, kind, case):
    """Euler on quads, p <= 3, runs four lanes per element (stage_kernel 2): residual and RK3
    steps equal the oracle's (FP64 1e-11, FP32 5e-5 / 2e-5) and the one-element-per-lane
    kernel's, on parallelograms (diagonal inverse mass), general quads (the element's FP64
    inverse) and meshes that mix both."""
    periodic, bcs = CASES[case]
    verts, quads = _quad_mesh(kind, 6, p)
    ora = R.Oracle(verts, quads, p, periodic=periodic)
    for tag, (typ, val) in bcs.items():
        if callable(val):
            ora.set_bc(tag, typ, fn=val)
        else:
            ora.set_bc(tag, typ, val)
    xy = ora.volume_points()
    u0 = ora.project(R.vortex_state(xy, xc=0.45, yc=0.55, beta=1.0, length=1.0))
    u0 = u0 + 1e-3 * np.random.default_rng(3 + p).standard_normal(u0.shape) * np.abs(u0).max(axis=(0, 1))
    r_ref = ora.residual(u0)
    c64 = dgx.Context(verts, quads, p, precision="fp64", periodic=periodic, bcs=bcs)
    c64.set_state(u0)
    dt = 0.5 * dgx.stable_dt(c64, 0.3)
    u_ref, _ = ora.rk3(u0, dt, 4)
    lane = _lane_context(verts, quads, p, precision="fp64", periodic=periodic, bcs=bcs)
    assert lane.info.stage_kernel == 0
    for prec, tr, ts in (("fp64", 1e-11, 1e-11), ("fp32", 5e-5, 2e-5)):
        c = c64 if prec == "fp64" else dgx.Context(verts, quads, p, precision=prec, periodic=periodic, bcs=bcs)
        assert c.info.stage_kernel == 2
        assert _rel(c.residual(u0), r_ref) < tr
        c.set_state(u0)
        c.rk3_steps(dt, 4)
        assert _rel(c.get_state(), u_ref) < ts
    assert _rel(c64.residual(u0), lane.residual(u0)) < 1e-13
    lane.set_state(u0)
    lane.rk3_steps(dt, 4)
    assert _rel(c64.get_state(), lane.get_state()) < 1e-13
