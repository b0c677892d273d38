"""Per-element geometry built on the device (dgx_get_geometry): bitwise equal to a numpy
restatement of the same operation sequence (affine inverse Jacobian and determinant, outward
unit normals, lengths, smallest edge; GeometricMapping::compute_mapping,
S/reference/mapping.cpp:22-50; DGSpace::compute_face_data, S/core/space.cpp:104-137), face codes
equal to the restated permutation rule (compute_face_permutation, :173-228), and normals and
lengths within 1e-14 of the oracle's independently computed face data.  Meshes: periodic and
bounded split-quad meshes and one with jittered interior vertices.
"""
import time

import numpy as np
import pytest

import dgx
import refcpu as R

pytestmark = pytest.mark.gpu


def _jittered(n, length, seed):
    verts, tris = R.split_quad_mesh(n, length)
    rng = np.random.default_rng(seed)
    interior = (verts[:, 0] > 0) & (verts[:, 0] < length) & (verts[:, 1] > 0) & (verts[:, 1] < length)
    verts = verts.copy()
    verts[interior] += rng.uniform(-0.2, 0.2, (interior.sum(), 2)) * (length / n)
    return verts, tris


def _face(V, f):
    a, b = f, (f + 1) % 3
    tx, ty = V[b, 0] - V[a, 0], V[b, 1] - V[a, 1]
    ln = np.sqrt(tx * tx + ty * ty)
    nx, ny = ty / ln, -tx / ln
    cx, cy = (V[0, 0] + V[1, 0] + V[2, 0]) / 3.0, (V[0, 1] + V[1, 1] + V[2, 1]) / 3.0
    if nx * (0.5 * (V[a, 0] + V[b, 0]) - cx) + ny * (0.5 * (V[a, 1] + V[b, 1]) - cy) < 0:
        nx, ny = -nx, -ny
    return nx, ny, ln


def _dist(p, q):
    dx, dy = p[0] - q[0], p[1] - q[1]
    return np.sqrt(dx * dx + dy * dy)


def _expected(verts, tris, nbr):
    K = len(tris)
    nbr = np.asarray(nbr).reshape(-1)
    geo = np.zeros((K, 14))
    hmin = np.zeros(K)
    codes = np.zeros(K, dtype=np.uint32)
    for e in range(K):
        V = verts[tris[e]]
        j00, j01, j10, j11 = V[1, 0] - V[0, 0], V[1, 1] - V[0, 1], V[2, 0] - V[0, 0], V[2, 1] - V[0, 1]
        det = j00 * j11 - j01 * j10
        inv = 1.0 / det
        geo[e, :5] = [j11 * inv, -j01 * inv, -j10 * inv, j00 * inv, abs(det)]
        hmin[e] = min(_dist(V[1], V[0]), _dist(V[2], V[1]), _dist(V[0], V[2]))
        for f in range(3):
            en, fn = nbr[(e * 3 + f) * 2], nbr[(e * 3 + f) * 2 + 1]
            nx, ny, ln = _face(V, f)
            if en < 0:
                c = 3 | 8
            else:
                isL = e < en or (e == en and f < fn)
                eL, fL, eR, fR = (e, f, en, fn) if isL else (en, fn, e, f)
                VL, VR = verts[tris[eL]], verts[tris[eR]]
                nLx, nLy, lnL = _face(VL, fL)
                a1, a2, b1, b2 = VL[fL], VL[(fL + 1) % 3], VR[fR], VR[(fR + 1) % 3]
                rev = _dist(a1, b2) < 1e-10 and _dist(a2, b1) < 1e-10
                same = _dist(a1, b1) < 1e-10 and _dist(a2, b2) < 1e-10
                if rev or same:
                    reverse = rev
                else:
                    t1, t2 = a2 - a1, b2 - b1
                    n1, n2 = np.sqrt(t1[0] * t1[0] + t1[1] * t1[1]), np.sqrt(t2[0] * t2[0] + t2[1] * t2[1])
                    reverse = (t1[0] / n1 * t2[0] / n2 + t1[1] / n1 * t2[1] / n2) < 0.0
                nx, ny, ln = (nLx, nLy, lnL) if isL else (-nLx, -nLy, lnL)
                c = fn | (4 if reverse else 0) | (8 if isL else 0)
            geo[e, 5 + 2 * f], geo[e, 6 + 2 * f], geo[e, 11 + f] = nx, ny, ln
            codes[e] |= np.uint32(c << (4 * f))
    return geo, hmin, codes


CASES = {
    "periodic": lambda: (*R.split_quad_mesh(5, 10.0), 3),
    "bounded": lambda: (*R.split_quad_mesh(4, 1.0), 0),
    "jittered_periodic": lambda: (*_jittered(6, 2.0, 3), 3),
    "jittered_half_periodic": lambda: (*_jittered(5, 1.0, 4), 1),
}


@pytest.mark.parametrize("case", list(CASES))
def test_device_geometry_is_bitwise_the_restated_formulas(case):
    verts, tris, periodic = CASES[case]()
    ctx = dgx.Context(verts, tris, 2, precision="fp64", periodic=periodic)
    geo, hmin, codes = ctx.geometry()
    eg, eh, ec = _expected(verts, tris, ctx._nbr)
    assert np.array_equal(codes, ec)
    assert np.array_equal(geo, eg), np.abs(geo - eg).max()
    assert np.array_equal(hmin, eh)
    # the oracle's own face data: outward normal and length of each element's face
    ora = R.Oracle(verts, tris, 2, basis="dubiner", periodic=periodic)
    for e in range(0, ctx.K, 3):
        for f in range(3):
            n, ln = ora.face_geom(e, f)
            # the R side carries the negated L normal: equal to its own outward normal to rounding
            assert np.abs(geo[e, 5 + 2 * f:7 + 2 * f] - n).max() < 1e-14
            assert abs(geo[e, 11 + f] - ln) < 1e-14 * ln


def test_singular_element_is_rejected():
    verts = np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [0.0, 1.0]])
    tris = np.array([[0, 1, 2], [0, 1, 3]], dtype=np.int32)
    with pytest.raises(dgx.DGXError, match="Singular Jacobian"):
        dgx.Context(verts, tris, 1, precision="fp64", periodic=0)


def test_geometry_build_at_two_million_elements():
    # 2 x 1024^2 elements: spot-check the device geometry against the restated formulas and
    # report the context creation time
    verts, tris = R.split_quad_mesh(1024, 10.0)
    t0 = time.perf_counter()
    ctx = dgx.Context(verts, tris, 3, precision="fp32")
    t1 = time.perf_counter()
    geo, hmin, codes = ctx.geometry()
    idx = np.random.default_rng(0).choice(ctx.K, 64, replace=False)
    sub_t = tris[idx]
    for k, e in enumerate(idx):
        V = verts[sub_t[k]]
        j00, j01, j10, j11 = V[1, 0] - V[0, 0], V[1, 1] - V[0, 1], V[2, 0] - V[0, 0], V[2, 1] - V[0, 1]
        det = j00 * j11 - j01 * j10
        assert geo[e, 4] == abs(det) and geo[e, 0] == j11 * (1.0 / det)
    # fully periodic: no face is a boundary face
    assert np.all(hmin > 0) and all(np.all(((codes >> (4 * f)) & 3) != 3) for f in range(3))
    print(f"context creation K={ctx.K}: {1e3 * (t1 - t0):.1f} ms")
