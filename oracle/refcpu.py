"""ctypes wrapper of the FP64 CPU oracle (oracle/refcpu.cpp).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker or the timed CPU baseline; never by the product.

Also holds the synthetic-input generators the checks share: the split-quad periodic mesh
and the isentropic vortex / Taylor-Green initial states, written with numpy from their
closed forms.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# the checker build (strict FP64: -ffp-contract=off, portable x86-64-v3) and the two timing
# builds of the CPU baseline, compiled like the reference's Release build on the host that
# runs them (-O3 -march=native, with and without -ffast-math; `make -C oracle timing`)
_LIB_PATHS = {"checker": os.path.join(_HERE, "_build", "librefcpu.so"),
              "native": os.path.join(_HERE, "_build", "librefcpu_native.so"),
              "fastmath": os.path.join(_HERE, "_build", "librefcpu_fastmath.so")}
_libs = {}

BC_FN = C.CFUNCTYPE(None, C.c_double, C.c_double, C.POINTER(C.c_double), C.c_void_p)


def build(target="all"):
    subprocess.check_call(["make", "-s", "-C", _HERE, target])


def lib(variant="checker"):
    if variant not in _libs:
        path = _LIB_PATHS[variant]
        if not os.path.exists(path):
            build("all" if variant == "checker" else "timing")
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.ref_create.restype = C.c_void_p
        L.ref_create.argtypes = [C.c_int, dp, C.c_int, ip, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                 C.c_int, C.c_int]
        L.ref_destroy.argtypes = [C.c_void_p]
        L.ref_last_error.restype = C.c_char_p
        L.ref_info.argtypes = [C.c_void_p, ip]
        L.ref_set_bc.argtypes = [C.c_void_p, C.c_char_p, C.c_int, dp, BC_FN, C.c_void_p]
        L.ref_set_euler_bc_points.argtypes = [C.c_void_p, C.c_int]
        L.ref_set_ns_average.argtypes = [C.c_void_p, C.c_int]
        L.ref_set_heat_flux.argtypes = [C.c_void_p, C.c_int]
        for name in ("ref_residual", "ref_rhs", "ref_project", "ref_eval_volume"):
            getattr(L, name).argtypes = [C.c_void_p, dp, dp]
        L.ref_apply_minv.argtypes = [C.c_void_p, dp]
        L.ref_volume_points.argtypes = [C.c_void_p, dp]
        L.ref_rk3.restype = C.c_double
        L.ref_rk3.argtypes = [C.c_void_p, dp, C.c_double, C.c_int]
        L.ref_rk3_threads.restype = C.c_double
        L.ref_rk3_threads.argtypes = [C.c_void_p, dp, C.c_double, C.c_int, C.c_int]
        L.ref_max_cfl.restype = C.c_double
        L.ref_max_cfl.argtypes = [C.c_void_p, dp, C.c_double]
        L.ref_density_range.argtypes = [C.c_void_p, dp, dp]
        L.ref_conservation_sums.argtypes = [C.c_void_p, dp, dp]
        L.ref_kinetic_energy.restype = C.c_double
        L.ref_kinetic_energy.argtypes = [C.c_void_p, dp]
        L.ref_quad_tri.argtypes = [C.c_int, dp, dp, dp]
        L.ref_gauss_legendre.argtypes = [C.c_int, dp, dp]
        L.ref_basis_eval.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, dp, dp]
        L.ref_mapping.argtypes = [C.c_int, dp, C.c_double, C.c_double, dp, dp, dp]
        L.ref_face_point.argtypes = [C.c_int, C.c_int, C.c_double, dp]
        L.ref_elem_data.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.ref_face_geom.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, dp]
        L.ref_interior_faces.argtypes = [C.c_void_p, ip]
        L.ref_boundary_faces.argtypes = [C.c_void_p, ip]
        L.ref_neighbors.argtypes = [C.c_void_p, ip]
        L.ref_volume_residual.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.ref_viscous_volume_residual.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.ref_interior_face_residual.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, dp, C.c_int]
        L.ref_boundary_face_residual.argtypes = [C.c_void_p, C.c_int, dp, dp, C.c_int]
        L.ref_rusanov.argtypes = [C.c_double, dp, dp, dp, dp]
        L.ref_viscous_flux.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.ref_penalty.restype = C.c_double
        L.ref_penalty.argtypes = [C.c_void_p, C.c_double]
        _libs[variant] = L
    return _libs[variant]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


# ----------------------------------------------------------------------------------------------
# Synthetic inputs
# ----------------------------------------------------------------------------------------------
def rect_quad_mesh(xmin, xmax, ymin, ymax, dx):
    """Structured quads on [xmin, xmax] x [ymin, ymax] as the reference's transfinite
    create_rectangular_mesh(dx, use_triangles=false) lays them out (S/utils/mesh_creation.cpp:
    212-243: nx = int((xmax - xmin)/dx), ny = int((ymax - ymin)/dx) divisions): vertex (i, j) at
    (xmin + i (xmax - xmin)/nx, ymin + j (ymax - ymin)/ny), quad (i, j) = (v_ij, v_i+1j, v_i+1j+1,
    v_ij+1) counter-clockwise, row-major (j outer)."""
    nx, ny = int((xmax - xmin) / dx), int((ymax - ymin) / dx)
    hx, hy = (xmax - xmin) / nx, (ymax - ymin) / ny
    X, Y = np.meshgrid(xmin + np.arange(nx + 1) * hx, ymin + np.arange(ny + 1) * hy, indexing="xy")
    verts = np.stack([X.ravel(), Y.ravel()], axis=1).astype(np.float64)
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()

    def vid(i, j):
        return j * (nx + 1) + i

    quads = np.stack([vid(ii, jj), vid(ii + 1, jj), vid(ii + 1, jj + 1), vid(ii, jj + 1)], axis=1).astype(np.int32)
    return verts, quads


def split_quad_mesh(n, length=10.0, x0=0.0, y0=0.0):
    """N x N squares on [x0, x0+L] x [y0, y0+L]; square (i,j) split along its (i,j)->(i+1,j+1)
    diagonal into T1 = (v_ij, v_i+1j, v_i+1j+1) and T2 = (v_ij, v_i+1j+1, v_ij+1), both CCW,
    ordered row-major (j outer, i inner), T1 then T2."""
    h = length / n
    idx = np.arange(n + 1)
    X, Y = np.meshgrid(x0 + idx * h, y0 + idx * h, indexing="xy")  # [j][i]
    verts = np.stack([X.ravel(), Y.ravel()], axis=1).astype(np.float64)

    def vid(i, j):
        return j * (n + 1) + i

    jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ii = ii.ravel()
    jj = jj.ravel()
    t1 = np.stack([vid(ii, jj), vid(ii + 1, jj), vid(ii + 1, jj + 1)], axis=1)
    t2 = np.stack([vid(ii, jj), vid(ii + 1, jj + 1), vid(ii, jj + 1)], axis=1)
    tris = np.empty((2 * n * n, 3), dtype=np.int32)
    tris[0::2] = t1
    tris[1::2] = t2
    return verts, tris


def splitmix64(x):
    """splitmix64 finaliser on uint64 arrays (Steele, Lea, Flood 2014)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_pm1(seed, k):
    """xi_e ~ U(-1,1) for element ids k: top 53 bits of splitmix64(seed ^ k)."""
    z = splitmix64(np.uint64(seed) ^ np.asarray(k, dtype=np.uint64))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0) * 2.0 - 1.0


def vortex_state(xy, beta=5.0, xc=5.0, yc=5.0, gamma=1.4, t=0.0, length=10.0, x0=0.0,
                 rho_pert=None):
    """Isentropic vortex, rho_inf = 1, u_inf = 1, v_inf = 0, p_inf = 1, advected to time t
    and wrapped periodically on [x0, x0+L].  Returns conserved U[..., 4]."""
    x = xy[..., 0] - t
    y = xy[..., 1]
    x = x0 + np.mod(x - x0, length)
    dx = x - xc
    dy = y - yc
    r2 = dx * dx + dy * dy
    f = np.exp(0.5 * (1.0 - r2))
    u = 1.0 - beta / (2 * np.pi) * f * dy
    v = beta / (2 * np.pi) * f * dx
    T = 1.0 - (gamma - 1.0) * beta * beta / (8 * gamma * np.pi ** 2) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    if rho_pert is not None:
        rho = rho * (1.0 + rho_pert)
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E], axis=-1)


def taylor_green_state(xy, gamma=1.4, mach=0.1, rho0=1.0, u0=1.0):
    x = xy[..., 0]
    y = xy[..., 1]
    c0 = u0 / mach
    p0 = rho0 * c0 * c0 / gamma
    u = u0 * np.sin(x) * np.cos(y)
    v = -u0 * np.cos(x) * np.sin(y)
    p = p0 + rho0 * u0 * u0 / 4.0 * (np.cos(2 * x) + np.cos(2 * y))
    rho = np.full_like(x, rho0)
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E], axis=-1)


# ----------------------------------------------------------------------------------------------
# Oracle context
# ----------------------------------------------------------------------------------------------
class Oracle:
    """FP64 CPU restatement of the reference's DGMesh + DGSpace + weak form + assembler +
    SSP-RK3 driver.  State layout [K][Np][4]."""

    def __init__(self, verts, tris, order, basis=None, vol_rule=0, periodic=3,
                 physics="euler", gamma=1.4, mu=1e-3, prandtl=0.72, sigma0=5.0,
                 ns_face_grad="reference", euler_bc_points="reference", legendre_endpoint="reference",
                 ns_average="reference", heat_flux="reference", build="checker"):
        L = self._L = lib(build)
        self._verts = np.ascontiguousarray(verts, dtype=np.float64)
        self._tris = np.ascontiguousarray(tris, dtype=np.int32)
        self.gamma = gamma
        nvpe = self._tris.shape[1]
        self.nvpe = nvpe
        # the reference's spaces: monomial basis on triangles, Legendre on quads
        basis = basis or ("legendre" if nvpe == 4 else "monomial")
        self._h = L.ref_create(len(self._verts), _dp(self._verts), len(self._tris), _ip(self._tris),
                               order, {"monomial": 0, "dubiner": 1, "legendre": 2}[basis], vol_rule, periodic,
                               {"euler": 0, "ns": 1}[physics], gamma, mu, prandtl, sigma0,
                               {"reference": 0, "correct": 1}[ns_face_grad], nvpe,
                               {"reference": 1, "correct": 0}[legendre_endpoint])
        if not self._h:
            raise RuntimeError(L.ref_last_error().decode())
        info = np.zeros(6, dtype=np.int32)
        L.ref_info(self._h, _ip(info))
        self.K, self.Np, self.Nq, self.Nfq, self.n_interior, self.n_boundary = map(int, info)
        self.order = order
        self._cbs = []
        # Q5: Euler boundary faces evaluate the condition at the reference's transposed read of
        # the face points ("reference") or at the points ("correct")
        L.ref_set_euler_bc_points(self._h, {"reference": 1, "correct": 0}[euler_bc_points])
        # Q4: viscous face fluxes as the reference forms them (interior half difference, entered
        # with the inviscid flux's sign) or the consistent +{Fv}.n
        L.ref_set_ns_average(self._h, {"reference": 1, "consistent": 0}[ns_average])
        # Q17: the energy viscous flux u.tau + q (q = -kappa grad T) as the reference forms it,
        # or u.tau - q (heat conducted down the temperature gradient)
        L.ref_set_heat_flux(self._h, {"reference": 1, "conduction": 0}[heat_flux])

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.ref_destroy(self._h)
            self._h = None

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(self._L.ref_last_error().decode())

    def set_bc(self, tag, bc_type, value=None, fn=None):
        t = {"far_field": 0, "slip_wall": 1, "no_slip_wall": 2}[bc_type]
        val = np.zeros(4) if value is None else np.ascontiguousarray(value, dtype=np.float64)
        cb = BC_FN(0)
        if fn is not None:
            def _cb(x, y, out, _user, _fn=fn):
                r = _fn(x, y)
                for k in range(4):
                    out[k] = float(r[k])
            cb = BC_FN(_cb)
        self._cbs.append((val, cb))
        self._L.ref_set_bc(self._h, tag.encode(), t, _dp(val), cb, None)

    def zeros(self):
        return np.zeros((self.K, self.Np, 4))

    def residual(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.zeros_like(u)
        self._chk(self._L.ref_residual(self._h, _dp(u), _dp(r)))
        return r

    def rhs(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        k = np.zeros_like(u)
        self._chk(self._L.ref_rhs(self._h, _dp(u), _dp(k)))
        return k

    def apply_minv(self, r):
        r = np.array(r, dtype=np.float64, order="C")
        self._chk(self._L.ref_apply_minv(self._h, _dp(r)))
        return r

    def rk3(self, u, dt, nsteps):
        """Returns (u_new, seconds in the stepping loop)."""
        u = np.array(u, dtype=np.float64, order="C")
        secs = self._L.ref_rk3(self._h, _dp(u), dt, nsteps)
        return u, secs

    def rk3_threads(self, u, dt, nsteps, nthreads):
        """rk3 on `nthreads` OpenMP threads (same result bit for bit): (u_new, seconds)."""
        u = np.array(u, dtype=np.float64, order="C")
        secs = self._L.ref_rk3_threads(self._h, _dp(u), dt, nsteps, nthreads)
        return u, secs

    def volume_points(self):
        xy = np.zeros((self.K, self.Nq, 2))
        self._chk(self._L.ref_volume_points(self._h, _dp(xy)))
        return xy

    def project(self, Uq):
        Uq = np.ascontiguousarray(Uq, dtype=np.float64)
        u = self.zeros()
        self._chk(self._L.ref_project(self._h, _dp(Uq), _dp(u)))
        return u

    def eval_volume(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        Uq = np.zeros((self.K, self.Nq, 4))
        self._chk(self._L.ref_eval_volume(self._h, _dp(u), _dp(Uq)))
        return Uq

    def max_cfl(self, u, dt):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return self._L.ref_max_cfl(self._h, _dp(u), dt)

    def density_range(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        o = np.zeros(2)
        self._L.ref_density_range(self._h, _dp(u), _dp(o))
        return o

    def conservation_sums(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        o = np.zeros(4)
        self._L.ref_conservation_sums(self._h, _dp(u), _dp(o))
        return o

    def kinetic_energy(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return self._L.ref_kinetic_energy(self._h, _dp(u))

    # element / face level
    def elem_data(self, e):
        j = np.zeros(self.Nq)
        d = np.zeros(self.Nq * self.Np * 2)
        self._L.ref_elem_data(self._h, e, _dp(j), _dp(d))
        return j, d.reshape(self.Nq, self.Np, 2)

    def face_geom(self, e, f):
        n = np.zeros(2)
        ln = np.zeros(1)
        self._L.ref_face_geom(self._h, e, f, _dp(n), _dp(ln))
        return n, float(ln[0])

    def interior_faces(self):
        o = np.zeros((max(self.n_interior, 1), 5), dtype=np.int32)
        self._L.ref_interior_faces(self._h, _ip(o))
        return o[: self.n_interior]

    def neighbors(self):
        o = np.zeros((self.K, self.nvpe, 2), dtype=np.int32)
        self._L.ref_neighbors(self._h, _ip(o))
        return o

    def boundary_faces(self):
        o = np.zeros((max(self.n_boundary, 1), 3), dtype=np.int32)
        self._L.ref_boundary_faces(self._h, _ip(o))
        return o[: self.n_boundary]

    def volume_residual(self, e, ue, viscous=False):
        ue = np.ascontiguousarray(ue, dtype=np.float64)
        R = np.zeros((self.Np, 4))
        fn = self._L.ref_viscous_volume_residual if viscous else self._L.ref_volume_residual
        self._chk(fn(self._h, e, _dp(ue), _dp(R)))
        return R

    def interior_face_residual(self, k, uL, uR, viscous=False):
        uL = np.ascontiguousarray(uL, dtype=np.float64)
        uR = np.ascontiguousarray(uR, dtype=np.float64)
        RL = np.zeros((self.Np, 4))
        RR = np.zeros((self.Np, 4))
        self._chk(self._L.ref_interior_face_residual(self._h, k, _dp(uL), _dp(uR), _dp(RL), _dp(RR),
                                                   int(viscous)))
        return RL, RR

    def boundary_face_residual(self, k, ue, viscous=False):
        ue = np.ascontiguousarray(ue, dtype=np.float64)
        R = np.zeros((self.Np, 4))
        self._chk(self._L.ref_boundary_face_residual(self._h, k, _dp(ue), _dp(R), int(viscous)))
        return R

    def viscous_flux(self, U, gradU):
        U = np.ascontiguousarray(U, dtype=np.float64)
        g = np.ascontiguousarray(gradU, dtype=np.float64).reshape(8)
        F = np.zeros(4)
        G = np.zeros(4)
        self._L.ref_viscous_flux(self._h, _dp(U), _dp(g), _dp(F), _dp(G))
        return F, G

    def penalty(self, h):
        return self._L.ref_penalty(self._h, h)


# ----------------------------------------------------------------------------------------------
# Tables
# ----------------------------------------------------------------------------------------------
def quad_tri(degree):
    n = lib().ref_quad_tri(degree, None, None, None)
    if n < 0:
        raise ValueError(lib().ref_last_error().decode())
    xi, eta, w = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().ref_quad_tri(degree, _dp(xi), _dp(eta), _dp(w))
    return xi, eta, w


def gauss_legendre(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().ref_gauss_legendre(n, _dp(x), _dp(w))
    return x, w


def basis_eval(kind, order, xi, eta, legendre_endpoint="reference"):
    np_ = (order + 1) ** 2 if kind == "legendre" else (order + 1) * (order + 2) // 2
    phi = np.zeros(np_)
    g = np.zeros(2 * np_)
    k = {"monomial": 0, "dubiner": 1, "legendre": 2}[kind] + (4 if legendre_endpoint == "correct" else 0)
    lib().ref_basis_eval(k, order, xi, eta, _dp(phi), _dp(g))
    return phi, g.reshape(np_, 2)


def mapping(verts, xi, eta):
    """verts [3][2] (triangle) or [4][2] (quad)."""
    V = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1)
    det = np.zeros(1)
    A = np.zeros(4)
    x = np.zeros(2)
    rc = lib().ref_mapping(len(V) // 2, _dp(V), xi, eta, _dp(det), _dp(A), _dp(x))
    if rc:
        raise RuntimeError(lib().ref_last_error().decode())
    return float(det[0]), A.reshape(2, 2), x


def face_point(f, s, nvpe=3):
    o = np.zeros(2)
    lib().ref_face_point(nvpe, f, s, _dp(o))
    return o


def rusanov(gamma, UL, UR, n):
    UL = np.ascontiguousarray(UL, dtype=np.float64)
    UR = np.ascontiguousarray(UR, dtype=np.float64)
    n = np.ascontiguousarray(n, dtype=np.float64)
    H = np.zeros(4)
    lib().ref_rusanov(gamma, _dp(UL), _dp(UR), _dp(n), _dp(H))
    return H
