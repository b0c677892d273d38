"""Ports of the reference's Navier-Stokes weak-form known-answer tests
(T/solver/navier_stokes_weak_formulation_test.cpp) against the oracle, on the reference's
two-triangle fixture (T/test_helpers.h:11-51) with the P1 shear-flow state of
make_shear_flow_setup (:123-177).  Expected values are rebuilt here with numpy from the
vertex coordinates, including the reference's read of the interleaved face-gradient storage
through a column-major Map (rows r -> (vec[r], vec[r + Nfq*Np])).
"""
import math

import numpy as np
import pytest

import refcpu as R

GAMMA = 1.4


def mesh2():
    verts = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    tris = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    return verts, tris


def prim2cons(W, g=GAMMA):
    rho, u, v, p = W
    return np.array([rho, rho * u, rho * v, p / (g - 1) + 0.5 * rho * (u * u + v * v)])


def shear_coeffs(verts, tris):
    # P1 monomial (1, eta, xi): c0 = f(v0), c1 = f(v2) - f(v0), c2 = f(v1) - f(v0)
    funcs = [lambda x, y: 1.0, lambda x, y: y, lambda x, y: 0.0,
             lambda x, y: 1.0 / (GAMMA - 1) + 0.5 * y * y]
    u = np.zeros((2, 3, 4))
    for e in range(2):
        V = verts[tris[e]]
        for v, f in enumerate(funcs):
            a = [f(*V[k]) for k in range(3)]
            u[e, :, v] = [a[0], a[2] - a[0], a[1] - a[0]]
    return u


def viscous_flux(U, g, mu, pr, gam=GAMMA):
    """Laminar NS fluxes from conserved U and its gradient g[v][d] (gas constant 1)."""
    rho, ru, rv, E = U
    u, v = ru / rho, rv / rho
    p = (gam - 1) * (E - 0.5 * rho * (u * u + v * v))
    gu = (g[1] - u * g[0]) / rho
    gv = (g[2] - v * g[0]) / rho
    gp = (gam - 1) * (g[3] - 0.5 * (u * u + v * v) * g[0] - rho * (u * gu + v * gv))
    gT = (rho * gp - p * g[0]) / (rho * rho)
    div = gu[0] + gv[1]
    txx = 2 * mu * gu[0] - 2 * mu / 3 * div
    tyy = 2 * mu * gv[1] - 2 * mu / 3 * div
    txy = mu * (gu[1] + gv[0])
    k = mu * gam / (pr * (gam - 1))
    q = -k * gT
    Fv = np.array([0, txx, txy, u * txx + v * txy + q[0]])
    Gv = np.array([0, txy, tyy, u * txy + v * tyy + q[1]])
    return Fv, Gv


def elem_geom(verts, tris, e):
    V = verts[tris[e]]
    JT = np.array([V[1] - V[0], V[2] - V[0]])
    A = np.linalg.inv(JT)                      # dxi_dx
    gref = np.array([[0.0, 0.0], [0.0, 1.0], [1.0, 0.0]])  # d(1, eta, xi)/d(xi, eta)
    return V, A, abs(np.linalg.det(JT)), gref @ A.T        # physical gradients [i][d]


def ref_point(V, x):
    return np.linalg.solve(np.array([V[1] - V[0], V[2] - V[0]]).T, x - V[0])


def face_points(V, f, s):
    a, b = V[f], V[(f + 1) % 3]
    return [(1 - 0.5 * (si + 1)) * a + 0.5 * (si + 1) * b for si in s]


def scrambled(gphys, nfq, np_=3):
    """Reference read of dphi_dx_face: storage vec[(q*Np+i)*2+d] = d-th physical derivative of
    phi_i at face point q (P1: independent of q); row r reads (vec[r], vec[r + nfq*Np])."""
    vec = np.array([gphys[(k // 2) % np_][k % 2] for k in range(2 * nfq * np_)])
    return np.array([[[vec[q * np_ + i], vec[q * np_ + i + nfq * np_]] for i in range(np_)] for q in range(nfq)])


def outward_normal(V, f):
    a, b = V[f], V[(f + 1) % 3]
    t = b - a
    n = np.array([t[1], -t[0]]) / np.linalg.norm(t)
    if n @ (0.5 * (a + b) - V.mean(axis=0)) < 0:
        n = -n
    return n, np.linalg.norm(t)


def make(mu, pr=0.72, sigma0=5.0, face_grad="reference", order=1):
    verts, tris = mesh2()
    return R.Oracle(verts, tris, order, basis="monomial", periodic=0, physics="ns", gamma=GAMMA, mu=mu,
                    prandtl=pr, sigma0=sigma0, ns_face_grad=face_grad), verts, tris


def test_penalty_parameter_scaling():
    # :185-200 (EXPECT_DOUBLE_EQ)
    o, _, _ = make(1e-2, 0.72, 7.5, order=2)
    assert o.penalty(0.125) == 7.5 * 1e-2 * 9 / 0.125
    assert o.penalty(1e-14) == 7.5 * 1e-2 * 9 / 1e-12
    with pytest.raises(RuntimeError):
        make(-1.0)


def test_uniform_flow_has_zero_viscous_residual():
    # :202-260
    o, verts, tris = make(1e-3)
    U = prim2cons([1.0, 0.5, 0.0, 1.0])
    for tag in ("Bottom", "Right", "Top", "Left"):
        o.set_bc(tag, "far_field", U)
    u = np.zeros((2, 3, 4))
    u[:, 0, :] = U
    for e in range(2):
        assert np.abs(o.volume_residual(e, u[e], viscous=True)).max() < 1e-12
    for k in range(o.n_interior):
        RL, RR = o.interior_face_residual(k, u[0], u[1], viscous=True)
        assert np.abs(RL).max() < 1e-12 and np.abs(RR).max() < 1e-12
    assert o.n_boundary == 4
    for k in range(o.n_boundary):
        assert np.abs(o.boundary_face_residual(k, u[o.boundary_faces()[k][0]], viscous=True)).max() < 1e-12


def test_shear_flow_volume_residual_matches_analytic():
    # :262-324
    mu, pr = 5e-2, 0.72
    o, verts, tris = make(mu, pr)
    u = shear_coeffs(verts, tris)
    xi, eta, w = R.quad_tri(2)
    kappa = mu * GAMMA / (pr * (GAMMA - 1))
    for e in range(2):
        V, A, detJ, gphys = elem_geom(verts, tris, e)
        grad_ru = gphys.T @ u[e, :, 1]
        grad_E = gphys.T @ u[e, :, 3]
        exp = np.zeros((3, 4))
        for q in range(3):
            wq = w[q] * detJ
            y = V[0][1] + (V[1][1] - V[0][1]) * xi[q] + (V[2][1] - V[0][1]) * eta[q]
            for i in range(3):
                exp[i, 1] -= wq * mu * gphys[i, 1]
                exp[i, 2] -= wq * mu * gphys[i, 0]
                gpy = (GAMMA - 1) * (grad_E[1] - y * grad_ru[1])
                exp[i, 3] -= wq * (mu * y - kappa * gpy) * gphys[i, 1]
        got = o.volume_residual(e, u[e], viscous=True)
        assert np.abs(got[:, 0]).max() < 1e-12
        assert np.allclose(got[:, 1:], exp[:, 1:], rtol=0, atol=1e-10)


def _interior_expected(o, verts, tris, u, mu, pr, read, average="reference"):
    e_l, f_l, e_r, f_r, rev = o.interior_faces()[0]
    s, wg = R.gauss_legendre(2)
    VL, AL, _, gL = elem_geom(verts, tris, e_l)
    VR, AR, _, gR = elem_geom(verts, tris, e_r)
    n, L = outward_normal(VL, f_l)
    sigma = o.penalty(L)
    pts = face_points(VL, f_l, s)
    GL = scrambled(gL, 2) if read == "reference" else np.array([gL, gL])
    GR = scrambled(gR, 2) if read == "reference" else np.array([gR, gR])
    RL, RR = np.zeros((3, 4)), np.zeros((3, 4))
    for q in range(2):
        qR = 1 - q if rev else q
        a, b = ref_point(VL, pts[q])
        c, d = ref_point(VR, pts[q])
        phL, phR = np.array([1, b, a]), np.array([1, d, c])
        UL, UR = phL @ u[e_l], phR @ u[e_r]
        gUL = np.einsum("id,iv->vd", GL[q], u[e_l])
        gUR = np.einsum("id,iv->vd", GR[qR], u[e_r])
        FvL, GvL = viscous_flux(UL, gUL, mu, pr)
        FvR, GvR = viscous_flux(UR, gUR, mu, pr)
        # the reference: each side along its own normal (a half difference); consistent: -{Fv}.n
        sL = 1.0 if average == "reference" else -1.0
        Fn = 0.5 * sL * (FvL * n[0] + GvL * n[1]) + 0.5 * (FvR * (-n[0]) + GvR * (-n[1])) - sigma * (UR - UL)
        RL -= wg[q] * L / 2 * np.outer(phL, Fn)
        RR += wg[q] * L / 2 * np.outer(phR, Fn)
    return RL, RR


def test_shear_flow_interior_face_matches_rederived_formula():
    # :326-405 -- pins the face-gradient read (quirk) and the half-difference average
    mu, pr = 5e-2, 0.72
    o, verts, tris = make(mu, pr)
    u = shear_coeffs(verts, tris)
    RL, RR = o.interior_face_residual(0, u[0], u[1], viscous=True)
    eL, eR = _interior_expected(o, verts, tris, u, mu, pr, "reference")
    assert np.allclose(RL, eL, rtol=0, atol=1e-10) and np.allclose(RR, eR, rtol=0, atol=1e-10)
    # the scrambled read makes this face term non-zero for a continuous linear field
    assert np.abs(RL).max() > 1e-3


def test_correct_face_gradients_cancel_for_continuous_linear_field():
    # with true gradients Fv_L = Fv_R and U_L = U_R on the face: Fn = 0
    mu, pr = 5e-2, 0.72
    o, verts, tris = make(mu, pr, face_grad="correct")
    u = shear_coeffs(verts, tris)
    RL, RR = o.interior_face_residual(0, u[0], u[1], viscous=True)
    assert np.abs(RL).max() < 1e-14 and np.abs(RR).max() < 1e-14
    eL, _ = _interior_expected(o, verts, tris, u, mu, pr, "correct")
    assert np.abs(eL).max() < 1e-14


def test_shear_flow_boundary_face_matches_rederived_formula():
    # :407-503 with the x-dependent far-field state (1, y, 0, 1) taken as conserved data
    mu, pr = 5e-2, 0.72
    o, verts, tris = make(mu, pr)
    bc = lambda x, y: prim2cons([1.0, y, 0.0, 1.0])  # noqa: E731
    for tag in ("Bottom", "Right", "Top", "Left"):
        o.set_bc(tag, "far_field", fn=bc)
    u = shear_coeffs(verts, tris)
    s, wg = R.gauss_legendre(2)
    for k, (e, f, has_bc) in enumerate(o.boundary_faces()):
        assert has_bc
        V, A, _, gp = elem_geom(verts, tris, e)
        n, L = outward_normal(V, f)
        sigma = o.penalty(L)
        Gs = scrambled(gp, 2)
        exp = np.zeros((3, 4))
        for q, x in enumerate(face_points(V, f, s)):
            a, b = ref_point(V, x)
            ph = np.array([1, b, a])
            UL = ph @ u[e]
            Fv, Gv = viscous_flux(UL, np.einsum("id,iv->vd", Gs[q], u[e]), mu, pr)
            Fn = (Fv * n[0] + Gv * n[1]) - sigma * (bc(*x) - UL)
            exp -= wg[q] * L / 2 * np.outer(ph, Fn)
        assert np.allclose(o.boundary_face_residual(k, u[e], viscous=True), exp, rtol=0, atol=1e-10)


def test_no_slip_boundary_generates_residual():
    # :505-536
    o, verts, tris = make(1e-3)
    for tag in ("Bottom", "Right", "Top", "Left"):
        o.set_bc(tag, "no_slip_wall", [1.0, 0.0, 0.0, 1.0])
    u = np.zeros((2, 3, 4))
    u[:, 0, :] = prim2cons([1.0, 0.5, 0.0, 1.0])
    assert np.linalg.norm(o.residual(u)) > 1e-6


def test_viscous_flux_closed_form_shear():
    # u = y, v = 0, rho = 1, p = 1: tau_xy = mu, no heat flux
    o, _, _ = make(0.03)
    y = 0.4
    U = prim2cons([1.0, y, 0.0, 1.0])
    g = np.zeros((4, 2))
    g[1] = [0.0, 1.0]
    g[3] = [0.0, y]  # dE/dy = u du/dy
    Fv, Gv = o.viscous_flux(U, g)
    assert np.allclose(Fv, [0, 0, 0.03, 0], atol=1e-15)
    assert np.allclose(Gv, [0, 0.03, 0, y * 0.03], atol=1e-15)


def _wall_state(verts, tris):
    """P1 monomial coefficients of a flow with non-zero normal and tangential velocity on every
    side of the unit square: rho = 1 + 0.2 x, u = 0.3 + 0.2 y, v = -0.4 + 0.1 x, p = 1 + 0.3 y
    (conserved variables taken linear between the vertex values)."""
    def W(x, y):
        return prim2cons([1.0 + 0.2 * x, 0.3 + 0.2 * y, -0.4 + 0.1 * x, 1.0 + 0.3 * y])
    u = np.zeros((2, 3, 4))
    for e in range(2):
        V = verts[tris[e]]
        a = [W(*V[k]) for k in range(3)]
        u[e] = np.array([a[0], a[2] - a[0], a[1] - a[0]])
    return u


def _cons2prim(U, gam=GAMMA):
    rho, ru, rv, E = U
    u, v = ru / rho, rv / rho
    return np.array([rho, u, v, (gam - 1) * (E - 0.5 * rho * (u * u + v * v))])


def _wall_expected(verts, tris, o, u, ghost):
    """Boundary residual rebuilt from the vertex coordinates: U_L at the GL face points,
    the ghost state from `ghost(W_L, n)`, Rusanov, R -= w |e|/2 phi H."""
    s, wg = R.gauss_legendre(2)
    out = []
    for k, (e, f, _) in enumerate(o.boundary_faces()):
        V, _, _, _ = elem_geom(verts, tris, e)
        n, L = outward_normal(V, f)
        exp = np.zeros((3, 4))
        for q, x in enumerate(face_points(V, f, s)):
            a, b = ref_point(V, x)
            ph = np.array([1, b, a])
            UL = ph @ u[e]
            UR = ghost(_cons2prim(UL), n)
            FL, sL = _euler_flux(UL, n)
            FR, sR = _euler_flux(UR, n)
            H = 0.5 * (FL + FR) - 0.5 * max(sL, sR) * (UR - UL)
            exp -= wg[q] * L / 2 * np.outer(ph, H)
        out.append((k, e, n, exp))
    return out


def test_euler_boundary_slip_wall_reflects_normal_velocity():
    # boundary_face_residual SLIP_WALL (euler_weak_formulation.cpp:214-224): the ghost state
    # keeps rho and p and mirrors the normal velocity, u_R = u_t - u_n.  The state has a
    # non-zero normal velocity on every wall, so a ghost that copied the interior state (or
    # dropped the reflection) gives a different flux; the reflected state carries no mass
    # through the wall even though u.n != 0
    verts, tris = mesh2()
    o = R.Oracle(verts, tris, 1, basis="monomial", periodic=0)
    for tag in ("Bottom", "Right", "Top", "Left"):
        o.set_bc(tag, "slip_wall")
    u = _wall_state(verts, tris)

    def reflect(W, n):
        un = W[1] * n[0] + W[2] * n[1]
        return prim2cons([W[0], W[1] - 2 * un * n[0], W[2] - 2 * un * n[1], W[3]])

    def copy(W, n):
        return prim2cons(W)

    for (k, e, n, exp), (_, _, _, wrong) in zip(_wall_expected(verts, tris, o, u, reflect),
                                                _wall_expected(verts, tris, o, u, copy)):
        got = o.boundary_face_residual(k, u[e])
        assert np.allclose(got, exp, rtol=0, atol=1e-13)
        assert np.abs(got[:, 0]).max() < 1e-14           # no mass through the wall
        assert np.abs(wrong[:, 0]).max() > 1e-2          # ... which an unreflected ghost has
        assert np.abs(got - wrong).max() > 1e-2


def test_euler_boundary_no_slip_ghost_is_the_primitive_condition():
    # boundary_face_residual NO_SLIP_WALL (euler_weak_formulation.cpp:225-227): the ghost state
    # is primitive_to_conserved(bc(x)), whatever the interior state
    verts, tris = mesh2()
    o = R.Oracle(verts, tris, 1, basis="monomial", periodic=0)
    Wb = [1.2, 0.0, 0.0, 0.9]
    for tag in ("Bottom", "Right", "Top", "Left"):
        o.set_bc(tag, "no_slip_wall", Wb)
    u = _wall_state(verts, tris)
    for k, e, n, exp in _wall_expected(verts, tris, o, u, lambda W, n: prim2cons(Wb)):
        got = o.boundary_face_residual(k, u[e])
        assert np.allclose(got, exp, rtol=0, atol=1e-13)
    # the same data read as a conserved state (the far-field branch) is a different flux
    wrong = _wall_expected(verts, tris, o, u, lambda W, n: np.array(Wb))
    k, e = wrong[0][0], wrong[0][1]
    assert np.abs(o.boundary_face_residual(k, u[e]) - wrong[0][3]).max() > 1e-2


def _euler_flux(U, n, gam=GAMMA):
    rho, ru, rv, E = U
    u, v = ru / rho, rv / rho
    p = (gam - 1) * (E - 0.5 * rho * (u * u + v * v))
    un = u * n[0] + v * n[1]
    return np.array([rho * un, ru * un + p * n[0], rv * un + p * n[1], (E + p) * un]), abs(un) + math.sqrt(gam * p / rho)


@pytest.mark.parametrize("points", ["reference", "correct"])
def test_euler_boundary_position_dependent_far_field(points):
    # boundary_face_residual FAR_FIELD with a function of position (euler_weak_formulation.cpp:
    # 188-213).  The face points are stored column-major, [x_0 .. x_{n-1}, y_0 .. y_{n-1}]
    # (space.cpp:141-164), and read back through a row-major n x 2 map, so point q is evaluated
    # at (flat[2q], flat[2q+1]): for two points (x_0, x_1) and (y_0, y_1).  "correct" reads (x_q, y_q).
    verts, tris = mesh2()
    o = R.Oracle(verts, tris, 1, basis="monomial", periodic=0, euler_bc_points=points)
    bc = lambda x, y: prim2cons([1.0 + 0.2 * x, 0.3 + y, 0.1 * x, 1.0 + 0.5 * y])  # noqa: E731
    for tag in ("Bottom", "Right", "Top", "Left"):
        o.set_bc(tag, "far_field", fn=bc)
    u = shear_coeffs(verts, tris)
    s, wg = R.gauss_legendre(2)
    seen_difference = False
    for k, (e, f, has_bc) in enumerate(o.boundary_faces()):
        V, _, _, _ = elem_geom(verts, tris, e)
        n, L = outward_normal(V, f)
        pts = face_points(V, f, s)
        flat = [pts[0][0], pts[1][0], pts[0][1], pts[1][1]]
        exp = np.zeros((3, 4))
        for q, x in enumerate(pts):
            xb = (flat[2 * q], flat[2 * q + 1]) if points == "reference" else tuple(x)
            seen_difference |= not np.allclose(xb, x)
            a, b = ref_point(V, x)
            ph = np.array([1, b, a])
            UL, UR = ph @ u[e], bc(*xb)
            FL, sL = _euler_flux(UL, n)
            FR, sR = _euler_flux(UR, n)
            H = 0.5 * (FL + FR) - 0.5 * max(sL, sR) * (UR - UL)
            exp -= wg[q] * L / 2 * np.outer(ph, H)
        assert np.allclose(o.boundary_face_residual(k, u[e]), exp, rtol=0, atol=1e-12)
    assert seen_difference == (points == "reference")


def test_consistent_viscous_face_flux_gives_the_physical_flux():
    # with true face gradients and the consistent flux, a continuous linear shear (u = y,
    # rho = p = 1: Fv = (0, 0, mu, 0), Gv = (0, mu, 0, y mu)) adds the physical flux Fv.n on the
    # interior face to M dU/dt, where the reference's half difference gives zero
    mu, pr = 5e-2, 0.72
    verts, tris = mesh2()
    o = R.Oracle(verts, tris, 1, basis="monomial", periodic=0, physics="ns", gamma=GAMMA, mu=mu,
                 prandtl=pr, sigma0=5.0, ns_face_grad="correct", ns_average="consistent")
    u = shear_coeffs(verts, tris)
    RL, RR = o.interior_face_residual(0, u[0], u[1], viscous=True)
    eL, eR = _interior_expected(o, verts, tris, u, mu, pr, "correct", average="consistent")
    assert np.allclose(RL, eL, rtol=0, atol=1e-14) and np.allclose(RR, eR, rtol=0, atol=1e-14)
    # momentum: the closed form Fn = (0, mu n_y, mu n_x, .) (the P1 energy carries a small heat
    # flux, so its component is left to the re-derived formula above)
    e_l, f_l, e_r, _, _ = o.interior_faces()[0]
    VL = verts[tris[e_l]]
    n, L = outward_normal(VL, f_l)
    s, wg = R.gauss_legendre(2)
    mom = np.zeros((3, 2))
    for q, x in enumerate(face_points(VL, f_l, s)):
        a, b = ref_point(VL, x)
        mom += wg[q] * L / 2 * np.outer([1, b, a], [mu * n[1], mu * n[0]])
    assert np.allclose(RL[:, 1:3], mom, rtol=0, atol=1e-14)
    assert np.abs(RL).max() > 1e-3


@pytest.mark.parametrize("heat", ["reference", "conduction"])
def test_heat_flux_sign(heat):
    # rest state with p = 1 + a x, rho = 1: grad T = (a, 0), q = -kappa grad T.  The reference's
    # energy flux is u.tau + q (navier_stokes_weak_formulation.cpp:76-77,91,95): heat up the
    # gradient; conduction is u.tau - q = kappa grad T
    mu, pr, a = 0.03, 0.72, 0.5
    verts, tris = mesh2()
    o = R.Oracle(verts, tris, 1, basis="monomial", periodic=0, physics="ns", gamma=GAMMA, mu=mu,
                 prandtl=pr, sigma0=5.0, heat_flux=heat)
    U = prim2cons([1.0, 0.0, 0.0, 1.0])
    g = np.zeros((4, 2))
    g[3] = [a / (GAMMA - 1), 0.0]  # dE/dx = dp/dx / (gamma - 1)
    Fv, Gv = o.viscous_flux(U, g)
    kappa = mu * GAMMA / (pr * (GAMMA - 1))
    assert Fv[3] == pytest.approx(-kappa * a if heat == "reference" else kappa * a, rel=1e-14)
    assert np.allclose(Fv[:3], 0, atol=1e-15) and np.allclose(Gv, 0, atol=1e-15)
