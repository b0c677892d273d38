"""The matrix-core tile kernels (FP32 Euler, p = 1 to 4, and Navier-Stokes, p = 3 and 4) on element
orders whose 64-element tiles are not compact patches of the mesh.

A context numbers its triangles in Morton order of the centroids, so a tile is a compact patch
with a few dozen faces that leave it.  With DGX_ELEMENT_ORDER=given the context keeps the
caller's order; a randomly shuffled element list then gives tiles of scattered elements, nearly
every face of which leaves the tile: more than 64 such faces per tile (the Euler kernel's second
and later passes over them) and more than 256 face points (the Navier-Stokes kernel's second and
later passes).  Both orders are held against the FP64 oracle on the shuffled mesh, with the
tolerances of the matrix-core kernel tests (5e-5 of the residual's largest entry), and against
each other.
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


def _context(given, *args, **kw):
    # DGX_ELEMENT_ORDER is read when the context is created
    if given:
        os.environ["DGX_ELEMENT_ORDER"] = "given"
    try:
        return dgx.Context(*args, **kw)
    finally:
        os.environ.pop("DGX_ELEMENT_ORDER", None)


def _shuffled_mesh(n, length, seed):
    verts, tris = R.split_quad_mesh(n, length)
    perm = np.random.default_rng(seed).permutation(len(tris))
    return verts, np.ascontiguousarray(tris[perm])


def _setup(physics, p, case, face_grad="reference", n=12):
    verts, tris = _shuffled_mesh(n, 2 * np.pi if case == "periodic" else 1.0, seed=p)
    okw = dict(physics="ns", mu=0.05, prandtl=0.71, sigma0=5.0, ns_face_grad=face_grad) if physics == "ns" else {}
    if case == "periodic":
        ora = R.Oracle(verts, tris, p, basis="dubiner", **okw)
        ckw = {}
        U = R.taylor_green_state(ora.volume_points())
    else:
        ora = R.Oracle(verts, tris, p, basis="dubiner", periodic=0, **okw)
        for tag, (typ, val) in MIXED.items():
            ora.set_bc(tag, typ, val)
        ckw = dict(periodic=0, bcs=MIXED)
        U = R.vortex_state(ora.volume_points(), xc=0.5, yc=0.5, beta=1.0, length=1.0)
    u0 = ora.project(U)
    rng = np.random.default_rng(p + 10)
    u0 = u0 + 1e-3 * rng.standard_normal(u0.shape) * np.abs(u0).max(axis=(0, 1))
    if physics == "ns":
        ckw["weak_form"] = dgx.NavierStokesWeakFormulation(1.4, 0.05, 0.71, 5.0, face_grad=face_grad)
    given = _context(True, verts, tris, p, basis="dubiner", precision="fp32", **ckw)
    morton = _context(False, verts, tris, p, basis="dubiner", precision="fp32", **ckw)
    return ora, given, morton, u0


@pytest.mark.parametrize("physics", ["euler", "ns"])
@pytest.mark.parametrize("p", [3, 4])
@pytest.mark.parametrize("case", ["periodic", "mixed"])
def test_scattered_tiles_residual_matches_oracle(physics, p, case):
    ora, given, morton, u0 = _setup(physics, p, case)
    assert given.info.stage_kernel == 1 and morton.info.stage_kernel == 1
    r_ref = ora.residual(u0)
    r_given, r_morton = given.residual(u0), morton.residual(u0)
    e_given, e_morton = _rel(r_given, r_ref), _rel(r_morton, r_ref)
    assert e_given < 5e-5 and e_morton < 5e-5, (e_given, e_morton)
    assert _rel(r_given, r_morton) < 1e-4


@pytest.mark.parametrize("p", [1, 2])
@pytest.mark.parametrize("case", ["periodic", "mixed"])
def test_scattered_tiles_low_order_euler(p, case):
    # At p = 1, 2 the 256 threads of a block take 2 or 3 points per face, so the first pass over
    # the external face points already reaches faces 64 and on, which are read from the tile's
    # full list rather than its fixed-stride copy of the first 64.
    ora, given, morton, u0 = _setup("euler", p, case)
    assert given.info.stage_kernel == 1 and morton.info.stage_kernel == 1
    r_ref = ora.residual(u0)
    r_given, r_morton = given.residual(u0), morton.residual(u0)
    e_given, e_morton = _rel(r_given, r_ref), _rel(r_morton, r_ref)
    assert e_given < 5e-5 and e_morton < 5e-5, (e_given, e_morton)
    assert _rel(r_given, r_morton) < 1e-4
    morton.set_state(u0)
    dt = 0.2 * dgx.stable_dt(morton, 0.3)
    for c in (given, morton):
        c.set_state(u0)
        c.rk3_steps(dt, 3)
    u_ref, _ = ora.rk3(u0, dt, 3)
    assert _rel(given.get_state(), u_ref) < 5e-5


@pytest.mark.parametrize("p", [3, 4])
def test_scattered_tiles_ns_plain_face_gradient(p):
    ora, given, morton, u0 = _setup("ns", p, "mixed", face_grad="correct")
    r_ref = ora.residual(u0)
    assert _rel(given.residual(u0), r_ref) < 5e-5
    assert _rel(morton.residual(u0), r_ref) < 5e-5


@pytest.mark.parametrize("physics", ["euler", "ns"])
def test_scattered_tiles_rk3_steps(physics):
    ora, given, morton, u0 = _setup(physics, 4, "periodic")
    morton.set_state(u0)
    dt = 0.2 * dgx.stable_dt(morton, 0.3)
    for c in (given, morton):
        c.set_state(u0)
        c.rk3_steps(dt, 3)
    u_ref, _ = ora.rk3(u0, dt, 3)
    assert _rel(given.get_state(), u_ref) < 5e-5
    assert _rel(morton.get_state(), u_ref) < 5e-5
