"""Legacy-VTK output (dgx_write_vtk, the reference's VTKWriter::write_euler_solution).

A state whose conserved variables are quadratic polynomials of (x, y) is represented exactly at
p = 2, so every point value in the file must equal the polynomial at the mapped subdivision
point (to the 10 significant digits the format carries); the header, counts, sub-triangle
connectivity and field order are checked against the format the reference writes.
"""
import numpy as np
import pytest

import dgx
import refcpu as R

pytestmark = pytest.mark.gpu

FIELDS = ["density", "u_velocity", "v_velocity", "velocity_magnitude", "pressure", "temperature", "mach",
          "rho", "rho_u", "rho_v", "E"]


def _poly(x, y):
    return np.stack([1.0 + 0.1 * x + 0.05 * y, 0.5 + 0.02 * x * x, 0.1 * y, 3.0 + 0.01 * x * y], axis=-1)


def _reference_subdivision(r):
    pts = [(i / r, j / r) for j in range(r + 1) for i in range(r + 1 - j)]

    def idx(i, j):
        return sum(r + 1 - jj for jj in range(j)) + i

    tris = []
    for j in range(r):
        for i in range(r - j):
            tris.append((idx(i, j), idx(i + 1, j), idx(i, j + 1)))
            if i < r - j - 1:
                tris.append((idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)))
    return np.array(pts), np.array(tris)


def _parse(path, nvc=3, cell_type="5"):
    lines = open(path).read().split("\n")
    assert lines[:4] == ["# vtk DataFile Version 3.0", "DGFEM Euler Solution", "ASCII", "DATASET UNSTRUCTURED_GRID"]
    k = 4
    assert lines[k].startswith("POINTS ") and lines[k].endswith(" double")
    n = int(lines[k].split()[1])
    xyz = np.array([[float(t) for t in lines[k + 1 + i].split()] for i in range(n)])
    assert all(lines[k + 1 + i].endswith(" 0.0") for i in range(n))
    k += 1 + n
    _, nc, size = lines[k].split()
    nc = int(nc)
    assert int(size) == (1 + nvc) * nc
    cells = np.array([[int(t) for t in lines[k + 1 + i].split()] for i in range(nc)])
    k += 1 + nc
    assert lines[k] == f"CELL_TYPES {nc}" and all(lines[k + 1 + i] == cell_type for i in range(nc))
    k += 1 + nc
    assert lines[k] == f"POINT_DATA {n}"
    k += 1
    data = {}
    for name in FIELDS:
        assert lines[k] == f"SCALARS {name} double 1" and lines[k + 1] == "LOOKUP_TABLE default"
        data[name] = np.array([float(t) for t in lines[k + 2:k + 2 + n]])
        k += 2 + n
    return xyz, cells, data


@pytest.mark.parametrize("basis", ["dubiner", "monomial"])
@pytest.mark.parametrize("r", [1, 2, 3])
def test_vtk_output_of_an_exact_quadratic_state(tmp_path, basis, r):
    verts, tris = R.split_quad_mesh(3, 1.0)
    ctx = dgx.Context(verts, tris, 2, basis=basis, precision="fp64", periodic=0)
    xy = ctx.volume_points()
    ctx.project(_poly(xy[..., 0], xy[..., 1]))
    ctx.write_vtk(str(tmp_path / "sol"), gamma=1.4, resolution_factor=r)
    xyz, cells, data = _parse(tmp_path / "sol.vtk")

    ref_pts, ref_tris = _reference_subdivision(r)
    npe = len(ref_pts)
    assert xyz.shape == (ctx.K * npe, 3) and cells.shape == (ctx.K * len(ref_tris), 4)
    # connectivity: the reference sub-triangles of element e, offset by e * npe
    want = np.concatenate([ref_tris + e * npe for e in range(ctx.K)])
    assert np.array_equal(cells[:, 0], np.full(len(cells), 3)) and np.array_equal(cells[:, 1:], want)
    # points: the affine map of each element
    V = verts[tris]  # [K][3][2]
    xi, eta = ref_pts[:, 0], ref_pts[:, 1]
    phys = (V[:, None, 0, :] * (1 - xi - eta)[None, :, None] + V[:, None, 1, :] * xi[None, :, None]
            + V[:, None, 2, :] * eta[None, :, None]).reshape(-1, 2)
    assert np.abs(xyz[:, :2] - phys).max() < 1e-9
    # fields
    U = _poly(phys[:, 0], phys[:, 1])
    rho, u, v = U[:, 0], U[:, 1] / U[:, 0], U[:, 2] / U[:, 0]
    p = 0.4 * (U[:, 3] - 0.5 * rho * (u * u + v * v))
    want = {"density": rho, "u_velocity": u, "v_velocity": v, "velocity_magnitude": np.hypot(u, v),
            "pressure": p, "temperature": p / rho, "mach": np.hypot(u, v) / np.sqrt(1.4 * p / rho),
            "rho": U[:, 0], "rho_u": U[:, 1], "rho_v": U[:, 2], "E": U[:, 3]}
    for name in FIELDS:
        err = np.abs(data[name] - want[name]).max() / np.abs(want[name]).max()
        assert err < 1e-9, (name, err)


@pytest.mark.parametrize("r", [1, 2, 4])
def test_vtk_output_of_quads(tmp_path, r):
    # rectangles (an affine map) at p = 2: the tensor basis holds the quadratic state exactly;
    # points (-1 + 2i/r, -1 + 2j/r) j outer, sub-quads (i,j),(i+1,j),(i+1,j+1),(i,j+1), VTK_QUAD
    verts, quads = dgx.rect_quad_mesh(0.0, 1.0, 0.0, 0.5, 0.25)
    ctx = dgx.Context(verts, quads, 2, precision="fp64", periodic=0)
    xy = ctx.volume_points()
    ctx.project(_poly(xy[..., 0], xy[..., 1]))
    ctx.write_vtk(str(tmp_path / "q"), resolution_factor=r)
    xyz, cells, data = _parse(tmp_path / "q.vtk", nvc=4, cell_type="9")
    n1 = r + 1
    npe = n1 * n1
    sub = np.array([[j * n1 + i, j * n1 + i + 1, (j + 1) * n1 + i + 1, (j + 1) * n1 + i]
                    for j in range(r) for i in range(r)])
    assert xyz.shape == (ctx.K * npe, 3) and cells.shape == (ctx.K * r * r, 5)
    assert np.array_equal(cells[:, 0], np.full(len(cells), 4))
    assert np.array_equal(cells[:, 1:], np.concatenate([sub + e * npe for e in range(ctx.K)]))
    V = verts[quads]  # [K][4][2], rectangles: x = x0 + (xi + 1)/2 (x1 - x0)
    s = -1.0 + 2.0 * np.arange(n1) / r
    xi, eta = np.meshgrid(s, s, indexing="xy")
    xi, eta = xi.ravel(), eta.ravel()
    phys = np.stack([V[:, None, 0, 0] + 0.5 * (xi + 1)[None] * (V[:, None, 1, 0] - V[:, None, 0, 0]),
                     V[:, None, 0, 1] + 0.5 * (eta + 1)[None] * (V[:, None, 3, 1] - V[:, None, 0, 1])],
                    axis=-1).reshape(-1, 2)
    assert np.abs(xyz[:, :2] - phys).max() < 1e-9
    U = _poly(phys[:, 0], phys[:, 1])
    assert np.abs(data["E"] - U[:, 3]).max() < 1e-9 * np.abs(U[:, 3]).max()
    assert np.abs(data["rho"] - U[:, 0]).max() < 1e-9 * np.abs(U[:, 0]).max()


def test_vtk_output_of_an_explicit_state_and_errors(tmp_path):
    verts, tris = R.split_quad_mesh(2, 1.0)
    ctx = dgx.Context(verts, tris, 2, precision="fp32", periodic=0)
    xy = ctx.volume_points()
    ctx.project(_poly(xy[..., 0], xy[..., 1]))
    u = ctx.get_state()
    ctx.set_state(np.zeros_like(u) + u * 0.5)  # the file must show u, not the context's state
    ctx.write_vtk(str(tmp_path / "explicit"), u=u, resolution_factor=2)
    xyz, _, data = _parse(tmp_path / "explicit.vtk")
    U = _poly(xyz[:, 0], xyz[:, 1])
    assert np.abs(data["E"] - U[:, 3]).max() < 1e-6 * np.abs(U[:, 3]).max()
    with pytest.raises(dgx.DGXError):
        ctx.write_vtk(str(tmp_path / "bad"), resolution_factor=0)
    with pytest.raises(dgx.DGXError):
        ctx.write_vtk(str(tmp_path / "missing_dir" / "x"))


# ---------------- relative L2 error norm on the device -----------------------------------------

def _l2_reference(Uq_num, Uq_ex_prim, verts, tris, p):
    # the reference tests' compute_variable_error restated in numpy: weights w_q |J|
    _, _, w = R.quad_tri(2 * p)
    V = verts[tris]
    det = np.abs((V[:, 1, 0] - V[:, 0, 0]) * (V[:, 2, 1] - V[:, 0, 1])
                 - (V[:, 2, 0] - V[:, 0, 0]) * (V[:, 1, 1] - V[:, 0, 1]))
    dw = w[None, :] * det[:, None]
    rho = Uq_num[..., 0]
    W = np.stack([rho, Uq_num[..., 1] / rho, Uq_num[..., 2] / rho,
                  0.4 * (Uq_num[..., 3] - 0.5 * (Uq_num[..., 1] ** 2 + Uq_num[..., 2] ** 2) / rho)], axis=-1)
    e2 = ((W - Uq_ex_prim) ** 2 * dw[..., None]).sum(axis=(0, 1))
    n2 = (Uq_ex_prim ** 2 * dw[..., None]).sum(axis=(0, 1))
    return np.where(n2 < 1e-14, np.sqrt(e2), np.sqrt(e2 / np.where(n2 < 1e-14, 1.0, n2)))


def _prim(U):
    rho = U[..., 0]
    return np.stack([rho, U[..., 1] / rho, U[..., 2] / rho,
                     0.4 * (U[..., 3] - 0.5 * (U[..., 1] ** 2 + U[..., 2] ** 2) / rho)], axis=-1)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_l2_error_matches_the_reference_formula(precision):
    verts, tris = R.split_quad_mesh(8)
    p = 3
    ora = R.Oracle(verts, tris, p, basis="dubiner")
    ctx = dgx.Context(verts, tris, p, basis="dubiner", precision=precision)
    xy = ora.volume_points()
    exact0 = R.vortex_state(xy)
    u0 = ora.project(exact0)
    ctx.set_state(u0)
    dt = dgx.stable_dt(ctx, 0.3)
    ctx.rk3_steps(dt, 10)
    # exact values at the context's own volume points (its point order is its own)
    exact = R.vortex_state(ctx.volume_points(), t=10 * dt)
    got = ctx.l2_error(exact)
    # the same norm from the oracle's evaluation of the GPU state at the oracle's points
    want = _l2_reference(ora.eval_volume(ctx.get_state()), _prim(R.vortex_state(xy, t=10 * dt)), verts, tris, p)
    assert np.all(want > 0)
    assert np.abs(got - want).max() < 1e-10 * want.max(), (got, want)
    assert np.allclose(ctx.l2_error(_prim(exact), primitive=True), got, rtol=1e-12, atol=0)


def test_l2_error_of_a_zero_field_is_absolute():
    # free stream with v = 0: the v field's exact norm is zero, so its error is the absolute root
    verts, tris = R.split_quad_mesh(4)
    ctx = dgx.Context(verts, tris, 2, precision="fp64")
    xy = ctx.volume_points()
    Uq = np.zeros(xy.shape[:2] + (4,))
    Uq[..., 0], Uq[..., 1], Uq[..., 3] = 1.0, 1.0, 1.0 / 0.4 + 0.5
    ctx.project(Uq)
    err = ctx.l2_error(Uq)
    assert np.all(err < 1e-12), err
    ex = Uq.copy()
    ex[..., 2] = 0.0
    ex[..., 1] = 1.1  # u off by 0.1 everywhere: relative error 0.1/1.1 in the u field
    err = ctx.l2_error(ex)
    assert abs(err[1] - 0.1 / 1.1) < 1e-12 and err[2] < 1e-12
