"""Pins the oracle's Euler residual, mass inverse and SSP-RK3 driver.

The reference has no value tests for the inviscid path (its Euler tests are commented out,
T/solver/weak_form_test.cpp:350-385), so the checks are invariants and independent
re-derivations: Rusanov consistency and antisymmetry, free-stream preservation, global
conservation, a from-scratch numpy evaluation of the volume and face terms on a two-element
mesh, basis independence of the discrete solution, exact monomial mass moments, the RK3
amplification polynomial, and convergence to the analytic vortex.
"""
import math

import numpy as np
import pytest

import refcpu as R

G = 1.4


def flux_F_G(U, g=G):
    rho, ru, rv, E = U
    u, v = ru / rho, rv / rho
    p = (g - 1) * (E - 0.5 * rho * (u * u + v * v))
    F = np.array([ru, ru * u + p, ru * v, u * (E + p)])
    Gf = np.array([rv, ru * v, rv * v + p, v * (E + p)])
    return F, Gf, p


def prim(rho, u, v, p, g=G):
    return np.array([rho, rho * u, rho * v, p / (g - 1) + 0.5 * rho * (u * u + v * v)])


UL = prim(1.1, 0.3, -0.2, 0.9)
UR = prim(0.8, -0.1, 0.4, 1.3)


def test_rusanov_consistency():
    n = np.array([0.6, 0.8])
    F, Gf, _ = flux_F_G(UL)
    assert np.allclose(R.rusanov(G, UL, UL, n), F * n[0] + Gf * n[1], rtol=1e-15, atol=1e-15)


def test_rusanov_antisymmetry_and_dissipation():
    # H(UL,UR,n) = -H(UR,UL,-n)  (the reference's commented-out test, weak_form_test.cpp:350-385)
    n = np.array([0.6, -0.8])
    H1 = R.rusanov(G, UL, UR, n)
    H2 = R.rusanov(G, UR, UL, -n)
    assert np.allclose(H1, -H2, rtol=0, atol=1e-15)
    # H = central - s/2 (UR - UL) with s = max(|u.n| + c)
    FL, GL, pL = flux_F_G(UL)
    FR, GR, pR = flux_F_G(UR)
    s = max(abs(UL[1] / UL[0] * n[0] + UL[2] / UL[0] * n[1]) + math.sqrt(G * pL / UL[0]),
            abs(UR[1] / UR[0] * n[0] + UR[2] / UR[0] * n[1]) + math.sqrt(G * pR / UR[0]))
    central = 0.5 * (FL * n[0] + GL * n[1] + FR * n[0] + GR * n[1])
    assert np.allclose(H1, central - 0.5 * s * (UR - UL), rtol=1e-14, atol=1e-14)


def two_triangle_mesh():
    # T/test_helpers.h:11-51
    verts = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    tris = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    return verts, tris


def random_state(K, Np, rng, scale=0.05):
    u = np.zeros((K, Np, 4))
    u[:, 0, :] = [1.0, 0.2, -0.1, 2.6]
    u[:, 1:, :] += scale * rng.standard_normal((K, Np - 1, 4))
    return u


@pytest.mark.parametrize("basis", ["monomial", "dubiner"])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_free_stream_preservation(basis, p):
    verts, tris = R.split_quad_mesh(4)
    o = R.Oracle(verts, tris, p, basis=basis)
    # constant state: monomial phi_0 = 1, Dubiner phi_0 = 2
    u = np.zeros((o.K, o.Np, 4))
    u[:, 0, :] = prim(1.2, 0.4, -0.7, 1.5) * (1.0 if basis == "monomial" else 0.5)
    assert np.abs(o.residual(u)).max() < 1e-12


@pytest.mark.parametrize("basis", ["monomial", "dubiner"])
def test_global_conservation_of_residual(basis):
    # periodic mesh: sum_e int phi_0 dU/dt = 0 for every variable
    rng = np.random.default_rng(1)
    verts, tris = R.split_quad_mesh(5)
    o = R.Oracle(verts, tris, 3, basis=basis)
    u = random_state(o.K, o.Np, rng)
    r = o.residual(u)
    # phi_0 is constant in both bases: its row has no volume term and the face terms cancel
    assert np.abs(r[:, 0, :].sum(axis=0)).max() < 1e-14 * np.abs(r[:, 0, :]).sum()
    assert np.abs(r[:, 0, :]).max() > 1e-3  # the individual rows are not trivially zero


def test_volume_and_face_terms_from_scratch_p1():
    """Two-triangle mesh, P1 monomial basis (1, eta, xi): rebuild the volume integral and the
    interior-face flux with numpy from the vertex coordinates and compare."""
    verts, tris = two_triangle_mesh()
    o = R.Oracle(verts, tris, 1, basis="monomial", periodic=0)
    rng = np.random.default_rng(3)
    u = random_state(2, 3, rng, 0.1)
    # degree-2 rule: (1/6,1/6), (1/6,2/3), (2/3,1/6) with weights 1/6
    qp = [(1 / 6, 1 / 6), (1 / 6, 2 / 3), (2 / 3, 1 / 6)]
    for e in range(2):
        V = verts[tris[e]]
        JT = np.array([V[1] - V[0], V[2] - V[0]])      # rows d x / d xi_a
        Ainv = np.linalg.inv(JT)                       # physical grad = Ainv @ ref grad
        detJ = abs(np.linalg.det(JT))
        Rref = np.zeros((3, 4))
        for (x, y) in qp:
            phi = np.array([1.0, y, x])
            gref = np.array([[0.0, 0.0], [0.0, 1.0], [1.0, 0.0]])
            gphys = gref @ Ainv.T
            U = phi @ u[e]
            F, Gf, _ = flux_F_G(U)
            Rref += (1 / 6) * detJ * (np.outer(gphys[:, 0], F) + np.outer(gphys[:, 1], Gf))
        assert np.allclose(o.volume_residual(e, u[e]), Rref, rtol=1e-14, atol=1e-14)
    # the diagonal is the only interior face: element 0 face 2 (v2->v0) vs element 1 face 0 (v0->v2)
    faces = o.interior_faces()
    assert faces.tolist() == [[0, 2, 1, 0, 1]]
    s, wg = R.gauss_legendre(2)
    n = np.array([-1.0, 1.0]) / math.sqrt(2)   # outward from element 0 (below the diagonal y = x)
    L = math.sqrt(2)
    RL, RR = np.zeros((3, 4)), np.zeros((3, 4))
    for q in range(2):
        t = 0.5 * (s[q] + 1)
        x = (1 - t) * np.array([1.0, 1.0]) + t * np.array([0.0, 0.0])   # physical point on face 2 of e0
        # reference coordinates of that point in each element
        def ref_coords(e):
            V = verts[tris[e]]
            return np.linalg.solve(np.array([V[1] - V[0], V[2] - V[0]]).T, x - V[0])
        a, b = ref_coords(0)
        c, d = ref_coords(1)
        phL, phR = np.array([1.0, b, a]), np.array([1.0, d, c])
        H = R.rusanov(G, phL @ u[0], phR @ u[1], n)
        RL -= wg[q] * L / 2 * np.outer(phL, H)
        RR += wg[q] * L / 2 * np.outer(phR, H)
    oL, oR = o.interior_face_residual(0, u[0], u[1])
    assert np.allclose(oL, RL, rtol=1e-13, atol=1e-14)
    assert np.allclose(oR, RR, rtol=1e-13, atol=1e-14)


def test_mass_inverse_from_exact_monomial_moments():
    # M_ref[i][j] = int x^(a_i+a_j) y^(b_i+b_j) = a! b! / (a+b+2)!; M_e = |J| M_ref
    p = 3
    verts, tris = R.split_quad_mesh(2)
    o = R.Oracle(verts, tris, p, basis="monomial")
    pw = [(i, d - i) for d in range(p + 1) for i in range(d + 1)]
    M = np.array([[math.factorial(a1 + a2) * math.factorial(b1 + b2) / math.factorial(a1 + a2 + b1 + b2 + 2)
                   for (a2, b2) in pw] for (a1, b1) in pw])
    detJ = (10 / 2) ** 2
    r = np.zeros((o.K, o.Np, 4))
    r[0, :, 0] = np.arange(1, o.Np + 1)
    k = o.apply_minv(r)
    assert np.allclose(k[0, :, 0], np.linalg.solve(detJ * M, r[0, :, 0]), rtol=1e-9)


def test_rk3_amplification_polynomial():
    """Around a uniform flow the stage operator is linear to O(eps^2): one SSP-RK3 step must
    act as 1 + z + z^2/2 + z^3/6 of the linearised operator."""
    verts, tris = R.split_quad_mesh(3)
    o = R.Oracle(verts, tris, 2, basis="dubiner")
    base = np.zeros((o.K, o.Np, 4))
    base[:, 0, :] = 0.5 * prim(1.0, 0.5, 0.25, 1.0)
    rng = np.random.default_rng(9)
    v = rng.standard_normal(base.shape)
    # at a uniform state every face is a tie of the Rusanov max(); the central difference is
    # still linear in v but carries an O(eps) term, hence the small eps
    eps = 1e-7
    dt = 0.02

    def J(w):
        s = np.abs(w).max()  # keep every perturbation eps-small
        return (o.rhs(base + eps * w / s) - o.rhs(base - eps * w / s)) * (s / (2 * eps))

    Jv = J(v)
    J2v = J(Jv)
    J3v = J(J2v)
    pred = v + dt * Jv + dt ** 2 / 2 * J2v + dt ** 3 / 6 * J3v
    up, _ = o.rk3(base + eps * v, dt, 1)
    um, _ = o.rk3(base - eps * v, dt, 1)
    step = (up - um) / (2 * eps)
    assert np.abs(step - pred).max() < 1e-6 * np.abs(pred).max()
    # a wrong stage weight moves the result far outside that band
    wrong = v + dt * Jv + dt ** 2 / 2 * J2v
    assert np.abs(step - wrong).max() > 1e3 * np.abs(step - pred).max()


def test_monomial_and_dubiner_give_the_same_solution():
    # same space, same quadrature: the discrete solutions agree up to roundoff
    verts, tris = R.split_quad_mesh(10)
    outs = []
    for basis in ("monomial", "dubiner"):
        o = R.Oracle(verts, tris, 2, basis=basis)
        xy = o.volume_points()
        u = o.project(R.vortex_state(xy))
        u, _ = o.rk3(u, 0.02, 10)
        outs.append(o.eval_volume(u))
    assert np.abs(outs[0] - outs[1]).max() < 1e-11


def test_vortex_converges_at_order_p_plus_one():
    errs = []
    p = 2
    for n in (16, 32):
        verts, tris = R.split_quad_mesh(n)
        o = R.Oracle(verts, tris, p, basis="dubiner")
        xy = o.volume_points()
        u = o.project(R.vortex_state(xy))
        T = 0.5
        steps = 4 * n
        u, _ = o.rk3(u, T / steps, steps)
        ex = R.vortex_state(xy, t=T)
        Uq = o.eval_volume(u)
        errs.append(np.linalg.norm(Uq[..., 0] - ex[..., 0]) / np.linalg.norm(ex[..., 0]))
    rate = math.log2(errs[0] / errs[1])
    assert rate > p + 0.5, (errs, rate)


def test_cfl_by_hand():
    # compute_max_cfl (S/solver/compressible_solver_base.cpp:114-170): dt max(|u|+c) / h_min
    verts, tris = R.split_quad_mesh(4)
    o = R.Oracle(verts, tris, 1, basis="monomial")
    U = prim(1.0, 3.0, 4.0, 1.0)
    u = np.zeros((o.K, o.Np, 4))
    u[:, 0, :] = U
    c = math.sqrt(G * 1.0 / 1.0)
    assert abs(o.max_cfl(u, 0.1) - 0.1 * (5.0 + c) / 2.5) < 1e-14
    u[3, 0, 0] = -1.0
    assert math.isinf(o.max_cfl(u, 0.1))
    lo, hi = o.density_range(u)
    assert lo == -1.0 and hi == 1.0


def test_connectivity_and_periodic_pairing():
    verts, tris = R.split_quad_mesh(4)
    o = R.Oracle(verts, tris, 1)
    nb = o.neighbors()
    assert (nb[..., 0] >= 0).all(), "fully periodic: no boundary faces"
    # symmetric pairing
    for e in range(o.K):
        for f in range(3):
            en, fn = nb[e, f]
            assert tuple(nb[en, fn]) == (e, f)
    # every interior face of this CCW mesh is traversed in opposite directions
    assert (o.interior_faces()[:, 4] == 1).all()
    assert o.n_interior == 3 * o.K // 2
    # outward unit normals
    for e in (0, 1, 5):
        V = verts[tris[e]]
        c = V.mean(axis=0)
        for f in range(3):
            n, L = o.face_geom(e, f)
            mid = 0.5 * (V[f] + V[(f + 1) % 3])
            assert abs(np.linalg.norm(n) - 1) < 1e-15 and n @ (mid - c) > 0
            assert abs(L - np.linalg.norm(V[(f + 1) % 3] - V[f])) < 1e-15


@pytest.mark.parametrize("physics", ["euler", "ns"])
def test_threaded_rk3_is_bitwise_the_sequential_one(physics):
    # the multi-core CPU baseline runs the same arithmetic in the same per-element order
    verts, tris = R.split_quad_mesh(6, 1.0)
    kw = dict(physics="ns", mu=0.01) if physics == "ns" else {}
    ora = R.Oracle(verts, tris, 2, basis="dubiner", periodic=1, **kw)  # Bottom/Top bounded
    ora.set_bc("Bottom", "far_field", [1.0, 1.0, 0.0, 1 / 0.4 + 0.5])
    ora.set_bc("Top", "no_slip_wall", [1.0, 0.0, 0.0, 1.0])
    xy = ora.volume_points()
    u0 = ora.project(R.vortex_state(xy, xc=0.5, yc=0.5, beta=1.0, length=1.0))
    a, _ = ora.rk3(u0, 1e-3, 3)
    b, _ = ora.rk3_threads(u0, 1e-3, 3, 4)
    assert np.array_equal(a, b)


def _linear_state_coeffs(verts, tris, p, funcs):
    """Monomial coefficients (1, eta, xi, ...) of a state whose four conserved variables are
    linear functions of (x, y): the P1 part from the vertex values, zero above."""
    K = len(tris)
    Np = (p + 1) * (p + 2) // 2
    u = np.zeros((K, Np, 4))
    for e in range(K):
        V = verts[tris[e]]
        for v, f in enumerate(funcs):
            a = [f(*V[k]) for k in range(3)]
            u[e, :3, v] = [a[0], a[2] - a[0], a[1] - a[0]]
    return u


L_BOX = 10.0
# rho = 2, m_x = 0.3 + 0.05 x, m_y = -0.2 + 0.08 y, E = 5 + 0.1 x - 0.04 y
LIN = [lambda x, y: 2.0, lambda x, y: 0.3 + 0.05 * x, lambda x, y: -0.2 + 0.08 * y,
       lambda x, y: 5.0 + 0.1 * x - 0.04 * y]


def _int_lin(a, b, c, L=L_BOX):
    """int over [0,L]^2 of a + b x + c y."""
    return a * L * L + (b + c) * L ** 3 / 2


def _int_sq(a, b, L=L_BOX):
    """int over [0,L]^2 of (a + b x)^2."""
    return L * (a * a * L + a * b * L * L + b * b * L ** 3 / 3)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_conservation_sums_closed_form(p):
    # int U dA over the periodic [0,10]^2 box for a constant and a linear state: a dropped
    # |J|, weight or variable moves the sums off the closed form
    verts, tris = R.split_quad_mesh(4, L_BOX)
    o = R.Oracle(verts, tris, p, basis="monomial")
    U = prim(1.3, 0.4, -0.7, 1.1)
    u = np.zeros((o.K, o.Np, 4))
    u[:, 0, :] = U
    assert np.allclose(o.conservation_sums(u), U * L_BOX ** 2, rtol=1e-14, atol=0)
    u = _linear_state_coeffs(verts, tris, p, LIN)
    exp = [_int_lin(2.0, 0, 0), _int_lin(0.3, 0.05, 0), _int_lin(-0.2, 0, 0.08), _int_lin(5.0, 0.1, -0.04)]
    assert np.allclose(o.conservation_sums(u), exp, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("basis", ["monomial", "dubiner"])
def test_kinetic_energy_closed_form(p, basis):
    # J = int (m_x^2 + m_y^2) / (2 rho) dA (the adjoint workload's functional).  Constant state:
    # 1/2 |m|^2 / rho times the area.  Linear momentum at constant density: a quadratic
    # integrand, integrated exactly by the degree-2p rule.  A missing 1/2, a dropped component
    # or a weight without |J| fails one of the two.
    verts, tris = R.split_quad_mesh(4, L_BOX)
    o = R.Oracle(verts, tris, p, basis=basis)
    rho, mx, my = 1.3, 0.4, -0.7
    u = np.zeros((o.K, o.Np, 4))
    u[:, 0, :] = np.array([rho, mx, my, 3.0]) * (1.0 if basis == "monomial" else 0.5)  # phi_0 = 1 / 2
    assert o.kinetic_energy(u) == pytest.approx(0.5 * (mx * mx + my * my) / rho * L_BOX ** 2, rel=1e-13)
    if basis == "monomial":
        u = _linear_state_coeffs(verts, tris, p, LIN)
    else:
        xy = o.volume_points()
        Uq = np.stack([f(xy[..., 0], xy[..., 1]) + 0 * xy[..., 0] for f in LIN], axis=-1)
        u = o.project(Uq)
    exp = (_int_sq(0.3, 0.05) + _int_sq(-0.2, 0.08)) / (2 * 2.0)
    assert o.kinetic_energy(u) == pytest.approx(exp, rel=1e-13)
