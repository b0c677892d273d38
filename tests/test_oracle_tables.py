"""Pins the oracle's quadrature rules, bases and reference mapping.

Known answers come from the reference's own tests (cited) and from closed forms; each check
is chosen so that a plausible slip (wrong ordering, dropped normalisation, transposed
Jacobian, wrong rule degree) fails it.
"""
import math

import numpy as np
import pytest

import refcpu as R


# ------------------------------------------------------------------ Gauss-Legendre
def test_gauss_legendre_two_point_kat():
    # T/quadrature/factory_test.cpp:11-27
    x, w = R.gauss_legendre(2)
    assert np.allclose(x, [-1 / math.sqrt(3), 1 / math.sqrt(3)], atol=1e-15)
    assert np.allclose(w, [1.0, 1.0], atol=1e-15)


@pytest.mark.parametrize("n", range(1, 9))
def test_gauss_legendre_exact_to_2n_minus_1(n):
    x, w = R.gauss_legendre(n)
    assert np.all(np.diff(x) > 0), "nodes ascending"
    for d in range(2 * n):
        exact = 0.0 if d % 2 else 2.0 / (d + 1)
        assert abs((w * x ** d).sum() - exact) < 1e-14
    # not exact at degree 2n (pins the point count)
    assert abs((w * x ** (2 * n)).sum() - 2.0 / (2 * n + 1)) > 1e-6


def test_gauss_legendre_three_point_closed_form():
    x, w = R.gauss_legendre(3)
    assert np.allclose(x, [-math.sqrt(0.6), 0.0, math.sqrt(0.6)], atol=1e-15)
    assert np.allclose(w, [5 / 9, 8 / 9, 5 / 9], atol=1e-15)


# ------------------------------------------------------------------ triangle rules
def _moment(a, b):
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)


@pytest.mark.parametrize("deg,npts", [(1, 1), (2, 3), (3, 4), (4, 6), (5, 7), (6, 12), (7, 13), (8, 16), (9, 19)])
def test_dunavant_degree_of_precision(deg, npts):
    # T/quadrature/factory_test.cpp:51-87 (weights sum to 1/2), :202-249 (degree of precision)
    xi, eta, w = R.quad_tri(deg)
    assert len(w) == npts
    # degrees 7 and 9 carry Dunavant's 15-digit constants; the others are exact to roundoff
    tol = 3e-15 if deg in (7, 9) else 1e-16
    assert abs(w.sum() - 0.5) < tol
    for d in range(deg + 1):
        for a in range(d + 1):
            assert abs((w * xi ** a * eta ** (d - a)).sum() - _moment(a, d - a)) < tol
    # some monomial of degree deg+1 is not integrated exactly
    errs = [abs((w * xi ** a * eta ** (deg + 1 - a)).sum() - _moment(a, deg + 1 - a)) for a in range(deg + 2)]
    assert max(errs) > 1e-8
    assert np.all(xi >= 0) and np.all(eta >= 0) and np.all(xi + eta <= 1)


def test_dunavant_published_points():
    # published Dunavant/Intrepid2 orbit values (40-digit tables)
    xi, eta, w = R.quad_tri(4)
    assert np.min(np.abs(xi - 0.44594849091596487577)) < 1e-15
    assert np.min(np.abs(w - 0.11169079483900573054)) < 1e-15
    xi, eta, w = R.quad_tri(6)
    assert np.min(np.abs(xi - 0.24928674517091042129)) < 1e-15
    assert np.min(np.abs(w - 0.025422453185103408460)) < 1e-15
    assert np.min(np.abs(xi - 0.053145049844816939902)) < 1e-15


def test_constant_integration():
    # T/quadrature/factory_test.cpp:338-361
    for deg in (2, 4, 6, 8):
        _, _, w = R.quad_tri(deg)
        assert abs(3.7 * w.sum() - 3.7 * 0.5) < 1e-14


# ------------------------------------------------------------------ monomial basis
def test_monomial_ordering_and_values():
    # 1, y, x, y^2, xy, x^2, ... (S/basis/monomial.cpp:34-40; T/basis/monomial_test.cpp:92,147-151)
    x, y = 0.3, 0.45
    phi, g = R.basis_eval("monomial", 2, x, y)
    assert np.allclose(phi, [1, y, x, y * y, x * y, x * x], atol=1e-16)
    assert np.allclose(g, [[0, 0], [0, 1], [1, 0], [0, 2 * y], [y, x], [2 * x, 0]], atol=1e-16)
    phi, g = R.basis_eval("monomial", 3, x, y)
    assert np.allclose(phi[6:], [y ** 3, x * y * y, x * x * y, x ** 3])
    assert np.allclose(g[7], [y * y, 2 * x * y])


# ------------------------------------------------------------------ Dubiner basis
def test_dubiner_closed_form_p1():
    # (i,j) order j outer, i inner: (0,0),(1,0),(0,1); phi = 2 sqrt((2i+1)(i+j+1)) ...
    x, y = 0.21, 0.37
    phi, g = R.basis_eval("dubiner", 1, x, y)
    assert np.allclose(phi, [2.0, 2 * math.sqrt(6) * (2 * x + y - 1), 2 * math.sqrt(2) * (3 * y - 1)], atol=1e-14)
    assert np.allclose(g, [[0, 0], [4 * math.sqrt(6), 2 * math.sqrt(6)], [0, 6 * math.sqrt(2)]], atol=1e-13)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_dubiner_gram_is_two_identity(p):
    # T/basis/dubiner_test.cpp:132-176: 0.5 * sum_q w_q phi phi^T = I
    xi, eta, w = R.quad_tri(2 * p)
    Phi = np.array([R.basis_eval("dubiner", p, a, b)[0] for a, b in zip(xi, eta)])
    G = 0.5 * (Phi.T * w) @ Phi
    assert np.abs(G - np.eye(len(G))).max() < 1e-13


@pytest.mark.parametrize("kind", ["dubiner", "monomial"])
@pytest.mark.parametrize("p", [2, 4])
def test_basis_gradient_matches_finite_differences(kind, p):
    # T/basis/dubiner_test.cpp:219-255
    h = 1e-6
    for x, y in [(0.2, 0.3), (0.6, 0.1), (0.05, 0.8)]:
        _, g = R.basis_eval(kind, p, x, y)
        fx = (R.basis_eval(kind, p, x + h, y)[0] - R.basis_eval(kind, p, x - h, y)[0]) / (2 * h)
        fy = (R.basis_eval(kind, p, x, y + h)[0] - R.basis_eval(kind, p, x, y - h)[0]) / (2 * h)
        assert np.abs(g[:, 0] - fx).max() < 1e-7 * max(1, np.abs(fx).max())
        assert np.abs(g[:, 1] - fy).max() < 1e-7 * max(1, np.abs(fy).max())


def test_dubiner_projection_reproduces_quadratic():
    # T/basis/dubiner_test.cpp:179-216
    p = 2
    xi, eta, w = R.quad_tri(2 * p)
    Phi = np.array([R.basis_eval("dubiner", p, a, b)[0] for a, b in zip(xi, eta)])
    f = lambda x, y: 1 + 2 * x - 3 * y + 0.5 * x * x + x * y - y * y  # noqa: E731
    c = 0.5 * (Phi.T * w) @ f(xi, eta)
    for x, y in [(0.1, 0.2), (0.3, 0.6), (0.7, 0.05)]:
        assert abs(R.basis_eval("dubiner", p, x, y)[0] @ c - f(x, y)) < 1e-13


# ------------------------------------------------------------------ mapping and faces
def test_mapping_det_and_inverse():
    # T/reference/mapping_test.cpp:98-126, :271-293 (J^T dxi_dx = I)
    V = np.array([[1.0, 1.0], [4.0, 2.0], [2.0, 5.0]])
    det, A, x = R.mapping(V, 0.25, 0.5)
    JT = np.array([[3.0, 1.0], [1.0, 4.0]])
    assert abs(det - 11.0) < 1e-14
    assert np.allclose(A, np.linalg.inv(JT), atol=1e-15)
    assert np.allclose(JT @ A, np.eye(2), atol=1e-15)
    assert np.allclose(x, [1 + 0.25 * 3 + 0.5 * 1, 1 + 0.25 * 1 + 0.5 * 4], atol=1e-15)
    with pytest.raises(RuntimeError):
        R.mapping(np.array([[0, 0], [1, 1], [2, 2.0]]), 0.2, 0.2)


def _face_integral(f, func, order=2):
    s, w = R.gauss_legendre(order + 1)
    a, b = [(0, 0), (1, 0), (0, 1)][f], [(0, 0), (1, 0), (0, 1)][(f + 1) % 3]
    L = math.hypot(b[0] - a[0], b[1] - a[1])
    tot = 0.0
    for si, wi in zip(s, w):
        xi = R.face_point(f, si)
        tot += func(xi[0], xi[1]) * wi * L / 2
    return tot


def test_face_integrals_of_monomials():
    # T/reference/face_integral_test.cpp:44-102,226-285
    assert abs(_face_integral(0, lambda x, y: 1.0) - 1.0) < 1e-14
    assert abs(_face_integral(1, lambda x, y: 1.0) - math.sqrt(2)) < 1e-14
    assert abs(_face_integral(1, lambda x, y: x) - math.sqrt(2) / 2) < 1e-14
    assert abs(_face_integral(1, lambda x, y: x * x * y * y, order=4) - math.sqrt(2) / 30) < 1e-14
    # face 2 runs (0,1) -> (0,0): t = (s+1)/2 maps to y = 1 - t
    assert np.allclose(R.face_point(2, -1.0), [0, 1]) and np.allclose(R.face_point(2, 1.0), [0, 0])
    assert np.allclose(R.face_point(1, -1.0), [1, 0]) and np.allclose(R.face_point(1, 1.0), [0, 1])
