"""Python binding of the dgx C ABI (include/dgx.h) over ctypes.

Mirrors the reference's compressible-solver interface for this path:
  * ``EulerWeakFormulation`` / ``NavierStokesWeakFormulation``  (physics parameters,
    I/weak_forms/euler_weak_formulation.hpp, navier_stokes_weak_formulation.hpp)
  * ``DGAssembler.assemble_euler_residual(u)``  -> M dU/dt      (S/solver/assembler.cpp:137-279)
  * ``EulerDGSolver`` / ``NavierStokesDGSolver``: apply_mass_inv, time_step_ssp_rk3,
    project_initial_condition, compute_max_cfl, compute_density_range, solve()
    (S/solver/compressible_solver_base.cpp)
State arrays use the reference layout [K][Np][4] in FP64.

The compute runs in libdgx.so (HIP, gfx950).  There is no CPU fallback: constructing a
context without the library or without a HIP device raises.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libdgx.so")

DGX_OK, DGX_E_INVALID, DGX_E_UNSUPPORTED, DGX_E_DEVICE, DGX_E_NONFINITE, DGX_E_NOMEM = range(6)
BASIS = {"monomial": 0, "dubiner": 1, "legendre": 2}
PRECISION = {"fp64": 0, "fp32": 1, "fp32_bf16vol": 2}
PHYSICS = {"euler": 0, "ns": 1}
BC_TYPES = {"far_field": 0, "slip_wall": 1, "no_slip_wall": 2, "periodic": 3, "none": -1}
QUIRK_NS_FACE_GRAD_REFERENCE = 1
QUIRK_EULER_BC_POINTS_REFERENCE = 2
QUIRK_LEGENDRE_ENDPOINT_REFERENCE = 4
QUIRK_NS_VISCOUS_AVERAGE_REFERENCE = 8
QUIRK_NS_HEAT_FLUX_REFERENCE = 16
BC_FN = C.CFUNCTYPE(None, C.c_double, C.c_double, C.POINTER(C.c_double), C.c_void_p)

# exported symbols, in the order include/dgx.h declares them
EXPORTS = [
    "dgx_mesh_split_quads", "dgx_mesh_connect", "dgx_mesh_connect_elements", "dgx_mesh_rect_quads",
    "dgx_mesh_read_gmsh", "dgx_create", "dgx_destroy", "dgx_last_error",
    "dgx_info", "dgx_max_elements", "dgx_set_state", "dgx_get_state", "dgx_volume_points", "dgx_project",
    "dgx_eval_volume", "dgx_get_geometry", "dgx_l2_error", "dgx_set_bc_function", "dgx_write_vtk",
    "dgx_residual", "dgx_apply_minv", "dgx_rk3_run", "dgx_rk3_steps", "dgx_rk3_steps_async",
    "dgx_synchronize", "dgx_rk3_profile", "dgx_conservation_sums", "dgx_conservation_sums_device", "dgx_max_cfl",
    "dgx_density_range", "dgx_kinetic_energy", "dgx_residual_vjp", "dgx_rhs_vjp",
    "dgx_kinetic_energy_grad", "dgx_rk3_vjp", "dgx_rk3_vjp_profile",
    "dgx_batch_create", "dgx_batch_destroy", "dgx_batch_instance", "dgx_batch_set_dt",
    "dgx_batch_rk3_steps", "dgx_batch_profile", "dgx_batch_get_tuning", "dgx_batch_set_tuning",
    "dgx_batch_autotune",
]


class DGXError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"dgx status {status}: {msg}")
        self.status = status


class MeshDesc(C.Structure):
    _fields_ = [("n_vertices", C.c_int32), ("vertices", C.POINTER(C.c_double)),
                ("n_elements", C.c_int32), ("elements", C.POINTER(C.c_int32)),
                ("face_neighbors", C.POINTER(C.c_int32)), ("n_bcs", C.c_int32),
                ("bc_types", C.POINTER(C.c_int32)), ("bc_values", C.POINTER(C.c_double)),
                ("vertices_per_element", C.c_int32)]


class SpaceDesc(C.Structure):
    _fields_ = [("order", C.c_int32), ("basis", C.c_int32), ("rule", C.c_int32)]


class Physics(C.Structure):
    _fields_ = [("kind", C.c_int32), ("gamma", C.c_double), ("mu", C.c_double),
                ("prandtl", C.c_double), ("sigma0", C.c_double), ("quirks", C.c_uint32)]


class BatchTuning(C.Structure):
    _fields_ = [("layout", C.c_int32), ("chunk", C.c_int32 * 4), ("blocks_per_cu", C.c_int32)]

    def as_dict(self):
        return {"layout": ["fused", "per_order", "concurrent"][self.layout], "chunk": list(self.chunk),
                "blocks_per_cu": self.blocks_per_cu}


class Info(C.Structure):
    _fields_ = [("n_elements", C.c_int32), ("n_modes", C.c_int32), ("n_volume_points", C.c_int32),
                ("n_face_points", C.c_int32), ("order", C.c_int32), ("basis", C.c_int32),
                ("precision", C.c_int32), ("physics", C.c_int32), ("n_boundary_faces", C.c_int32),
                ("device", C.c_int32), ("bytes_per_stage_alg", C.c_double), ("stage_kernel", C.c_int32)]


_lib = None


def load():
    """Load libdgx.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DGXError(DGX_E_DEVICE, f"{LIB_PATH} is missing: build it (python __graft_entry__.py or make)")
    L = C.CDLL(LIB_PATH)
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int32)
    vp = C.c_void_p
    L.dgx_mesh_split_quads.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_double, dp, ip]
    L.dgx_mesh_connect.argtypes = [C.c_int32, dp, C.c_int32, ip, C.c_int32, ip]
    L.dgx_mesh_connect_elements.argtypes = [C.c_int32, dp, C.c_int32, ip, C.c_int32, C.c_int32, ip]
    L.dgx_mesh_rect_quads.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, dp, ip]
    i32p = C.POINTER(C.c_int32)
    L.dgx_mesh_read_gmsh.argtypes = [C.c_char_p, i32p, i32p, i32p, i32p, dp, ip, ip, C.c_char_p]
    L.dgx_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(SpaceDesc), C.POINTER(Physics), C.c_int32,
                             C.POINTER(vp)]
    L.dgx_destroy.argtypes = [vp]
    L.dgx_destroy.restype = None
    L.dgx_last_error.restype = C.c_char_p
    L.dgx_info.argtypes = [vp, C.POINTER(Info)]
    for n in ("dgx_set_state", "dgx_get_state", "dgx_volume_points", "dgx_project", "dgx_eval_volume",
              "dgx_apply_minv", "dgx_conservation_sums", "dgx_density_range", "dgx_kinetic_energy"):
        getattr(L, n).argtypes = [vp, dp]
    L.dgx_l2_error.argtypes = [vp, dp, C.c_int32, dp]
    L.dgx_set_bc_function.argtypes = [vp, C.c_int32, BC_FN, vp]
    L.dgx_get_geometry.argtypes = [vp, dp, dp, C.POINTER(C.c_uint32)]
    L.dgx_write_vtk.argtypes = [vp, dp, C.c_char_p, C.c_double, C.c_int32]
    L.dgx_residual.argtypes = [vp, dp, dp]
    L.dgx_conservation_sums_device.argtypes = [vp, vp]
    L.dgx_max_elements.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_int32)]
    L.dgx_rk3_run.argtypes = [vp, C.c_double, C.c_int32, C.POINTER(C.c_int32)]
    L.dgx_rk3_steps.argtypes = [vp, C.c_double, C.c_int32]
    L.dgx_rk3_steps_async.argtypes = [vp, C.c_double, C.c_int32]
    L.dgx_synchronize.argtypes = [vp]
    L.dgx_rk3_profile.argtypes = [vp, C.c_double, C.c_int32, dp, dp]
    L.dgx_max_cfl.argtypes = [vp, C.c_double, dp]
    L.dgx_residual_vjp.argtypes = [vp, dp, dp, dp]
    L.dgx_rhs_vjp.argtypes = [vp, dp, dp, dp]
    L.dgx_kinetic_energy_grad.argtypes = [vp, dp, dp]
    L.dgx_rk3_vjp.argtypes = [vp, C.c_double, C.c_int32, C.c_int32, dp, dp]
    L.dgx_rk3_vjp_profile.argtypes = [vp, C.c_double, C.c_int32, C.c_int32, dp, dp, dp]
    L.dgx_batch_create.argtypes = [C.c_int32, C.POINTER(MeshDesc), C.POINTER(SpaceDesc), C.POINTER(Physics),
                                   C.c_int32, C.POINTER(vp)]
    L.dgx_batch_destroy.argtypes = [vp]
    L.dgx_batch_destroy.restype = None
    L.dgx_batch_instance.argtypes = [vp, C.c_int32, C.POINTER(vp)]
    L.dgx_batch_set_dt.argtypes = [vp, dp]
    L.dgx_batch_rk3_steps.argtypes = [vp, C.c_int32]
    L.dgx_batch_profile.argtypes = [vp, C.c_int32, C.c_int32, dp]
    L.dgx_batch_get_tuning.argtypes = [vp, C.POINTER(BatchTuning)]
    L.dgx_batch_set_tuning.argtypes = [vp, C.POINTER(BatchTuning)]
    L.dgx_batch_autotune.argtypes = [vp, C.c_int32, C.POINTER(BatchTuning)]
    for n in EXPORTS:
        f = getattr(L, n)
        if n not in ("dgx_destroy", "dgx_last_error", "dgx_batch_destroy"):
            f.restype = C.c_int32
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _check(st):
    if st != DGX_OK:
        raise DGXError(st, load().dgx_last_error().decode())


# ----------------------------------------------------------------------------------------------
# mesh helpers
# ----------------------------------------------------------------------------------------------
def split_quad_mesh(n, length=10.0, x0=0.0, y0=0.0):
    verts = np.zeros(((n + 1) * (n + 1), 2))
    tris = np.zeros((2 * n * n, 3), dtype=np.int32)
    _check(load().dgx_mesh_split_quads(n, length, x0, y0, _dp(verts), _ip(tris)))
    return verts, tris


def rect_quad_mesh(xmin, xmax, ymin, ymax, dx):
    """Structured quads as MeshCreator::create_rectangular_mesh(dx, use_triangles=false) lays
    them out: int((xmax - xmin)/dx) x int((ymax - ymin)/dx) cells."""
    nx, ny = int((xmax - xmin) / dx), int((ymax - ymin) / dx)
    verts = np.zeros(((nx + 1) * (ny + 1), 2))
    quads = np.zeros((nx * ny, 4), dtype=np.int32)
    _check(load().dgx_mesh_rect_quads(nx, ny, xmin, xmax, ymin, ymax, _dp(verts), _ip(quads)))
    return verts, quads


def read_gmsh(path):
    """(vertices, elements, face neighbours, boundary tag names) of a GMSH ASCII mesh; boundary
    faces carry the index of their physical group in the names (pass tags=names to Context)."""
    L = load()
    n = [C.c_int32() for _ in range(4)]
    p = str(path).encode()
    _check(L.dgx_mesh_read_gmsh(p, *[C.byref(x) for x in n], None, None, None, None))
    nv, K, nvpe, nt = (x.value for x in n)
    verts = np.zeros((nv, 2))
    elems = np.zeros((K, nvpe), dtype=np.int32)
    nbr = np.zeros((K, nvpe, 2), dtype=np.int32)
    names = C.create_string_buffer(64 * max(nt, 1))
    _check(L.dgx_mesh_read_gmsh(p, *[C.byref(x) for x in n], _dp(verts), _ip(elems), _ip(nbr), names))
    tags = [names.raw[64 * k:64 * (k + 1)].split(b"\0")[0].decode() for k in range(nt)]
    return verts, elems, nbr, tags


def connect(verts, elems, periodic=3):
    """Face neighbours [K][nv][2] of triangles (nv = 3) or quadrilaterals (nv = 4)."""
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    elems = np.ascontiguousarray(elems, dtype=np.int32)
    nv = elems.shape[1]
    nbr = np.zeros((len(elems), nv, 2), dtype=np.int32)
    _check(load().dgx_mesh_connect_elements(len(verts), _dp(verts), len(elems), _ip(elems), nv, periodic, _ip(nbr)))
    return nbr


# ----------------------------------------------------------------------------------------------
# weak forms (physics parameters)
# ----------------------------------------------------------------------------------------------
class EulerWeakFormulation:
    """EulerWeakFormulation(gamma) (I/weak_forms/euler_weak_formulation.hpp)."""
    kind = "euler"

    def __init__(self, gamma=1.4):
        self.gamma = gamma
        self.mu, self.prandtl, self.sigma0, self.quirks = 0.0, 0.72, 5.0, 0

    def has_viscous_terms(self):
        return False


class NavierStokesWeakFormulation(EulerWeakFormulation):
    """NavierStokesWeakFormulation(gamma, mu, Pr, sigma0) with the reference defaults
    mu = 1e-3, Pr = 0.72, sigma0 = 5 (I/solver/dg_solver.hpp:283-285).  ``face_grad`` selects
    the reference's face-gradient read ("reference") or the mathematically correct one."""
    kind = "ns"

    def __init__(self, gamma=1.4, dynamic_viscosity=1e-3, prandtl=0.72, penalty_prefactor=5.0,
                 face_grad="reference", viscous_average="reference", heat_flux="reference"):
        """face_grad: the reference's read of the face gradients or the "correct" one;
        viscous_average: the viscous face fluxes as the reference forms them (interior half
        difference, subtracted like the inviscid flux) or the "consistent" +{Fv}.n; heat_flux: the energy flux u.tau + q as the reference forms it
        (heat up the temperature gradient) or "conduction" (u.tau - q)."""
        super().__init__(gamma)
        if dynamic_viscosity <= 0:
            raise ValueError("Dynamic viscosity must be positive for Navier-Stokes weak formulation")
        if prandtl <= 0:
            raise ValueError("Prandtl number must be positive for Navier-Stokes weak formulation")
        if penalty_prefactor <= 0:
            raise ValueError("Penalty prefactor must be positive for Navier-Stokes weak formulation")
        self.mu, self.prandtl, self.sigma0 = dynamic_viscosity, prandtl, penalty_prefactor
        if (face_grad not in ("reference", "correct") or viscous_average not in ("reference", "consistent")
                or heat_flux not in ("reference", "conduction")):
            raise ValueError("face_grad: reference|correct, viscous_average: reference|consistent, "
                             "heat_flux: reference|conduction")
        self.quirks = ((QUIRK_NS_FACE_GRAD_REFERENCE if face_grad == "reference" else 0)
                       | (QUIRK_NS_VISCOUS_AVERAGE_REFERENCE if viscous_average == "reference" else 0)
                       | (QUIRK_NS_HEAT_FLUX_REFERENCE if heat_flux == "reference" else 0))

    def compute_penalty_parameter(self, p, h):
        return self.sigma0 * self.mu * (p + 1) * (p + 1) / max(h, 1e-12)

    def has_viscous_terms(self):
        return True


# ----------------------------------------------------------------------------------------------
# context
# ----------------------------------------------------------------------------------------------
class Context:
    """One mesh + space + physics + precision on the GPU.  Layout of every host state array:
    [K][Np][4] FP64."""

    def __init__(self, verts, tris, order, weak_form=None, basis=None, precision="fp32",
                 rule=0, periodic=3, neighbors=None, bcs=None, euler_bc_points="reference",
                 legendre_endpoint="reference", tags=None):
        """tris: [K][3] triangles or [K][4] quadrilaterals (counter-clockwise).  basis: dubiner
        (default) or monomial on triangles, legendre on quadrilaterals.
        bcs: {tag: (type, value)} with tag in `tags` (default Bottom/Right/Top/Left, the boundary
        indices of `connect`; read_gmsh returns a mesh's own names) and value a constant state
        (FAR_FIELD conserved, NO_SLIP primitive) or a function (x, y) -> state.
        euler_bc_points: where Euler boundary faces evaluate a position-dependent condition:
        "reference" (the reference's transposed read of the face points) or "correct".
        legendre_endpoint: the Legendre derivative at +1 as the reference computes it
        ("reference") or P'_n(1) = n(n+1)/2 ("correct")."""
        if euler_bc_points not in ("reference", "correct"):
            raise ValueError("euler_bc_points must be 'reference' or 'correct'")
        if legendre_endpoint not in ("reference", "correct"):
            raise ValueError("legendre_endpoint must be 'reference' or 'correct'")
        L = load()
        wf = weak_form or EulerWeakFormulation()
        self.weak_form = wf
        self.order = order
        self._verts = np.ascontiguousarray(verts, dtype=np.float64)
        self._tris = np.ascontiguousarray(tris, dtype=np.int32)
        self._nbr = (connect(self._verts, self._tris, periodic) if neighbors is None
                     else np.ascontiguousarray(neighbors, dtype=np.int32))
        # boundary-condition table indexed by tag 0..3 = Bottom, Right, Top, Left
        bcs = bcs or {}
        tags = list(tags) if tags is not None else ["Bottom", "Right", "Top", "Left"]
        bc_t = np.full(max(len(tags), 1), -1, dtype=np.int32)
        bc_v = np.zeros((max(len(tags), 1), 4))
        bc_fns = {}
        for name, (typ, val) in bcs.items():
            k = tags.index(name)
            bc_t[k] = BC_TYPES[typ]
            if callable(val):
                bc_fns[k] = val
            else:
                bc_v[k] = val
        self._bc_t, self._bc_v = bc_t, bc_v
        nvpe = self._tris.shape[1]
        mesh = MeshDesc(len(self._verts), _dp(self._verts), len(self._tris), _ip(self._tris),
                        _ip(self._nbr), len(tags), _ip(bc_t), _dp(bc_v), nvpe)
        basis = basis or ("legendre" if nvpe == 4 else "dubiner")
        space = SpaceDesc(order, BASIS[basis], rule)
        quirks = (wf.quirks | (QUIRK_EULER_BC_POINTS_REFERENCE if euler_bc_points == "reference" else 0)
                  | (QUIRK_LEGENDRE_ENDPOINT_REFERENCE if legendre_endpoint == "reference" else 0))
        phys = Physics(PHYSICS[wf.kind], wf.gamma, wf.mu, wf.prandtl, wf.sigma0, quirks)
        h = C.c_void_p()
        _check(L.dgx_create(C.byref(mesh), C.byref(space), C.byref(phys), PRECISION[precision], C.byref(h)))
        self._h = h
        info = Info()
        _check(L.dgx_info(h, C.byref(info)))
        self.info = info
        self.K, self.Np, self.Nq, self.Nfq = info.n_elements, info.n_modes, info.n_volume_points, info.n_face_points
        self.precision = precision
        self._bc_cbs = []
        for k, fn in bc_fns.items():
            def _cb(x, y, out, _user, _fn=fn):
                r = _fn(x, y)
                for v in range(4):
                    out[v] = float(r[v])
            cb = BC_FN(_cb)
            self._bc_cbs.append(cb)  # called again whenever a later condition is set
            _check(L.dgx_set_bc_function(h, k, cb, None))

    def close(self):
        if getattr(self, "_h", None):
            load().dgx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # state
    def set_state(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.K, self.Np, 4)
        _check(load().dgx_set_state(self._h, _dp(u)))

    def get_state(self):
        u = np.zeros((self.K, self.Np, 4))
        _check(load().dgx_get_state(self._h, _dp(u)))
        return u

    def volume_points(self):
        xy = np.zeros((self.K, self.Nq, 2))
        _check(load().dgx_volume_points(self._h, _dp(xy)))
        return xy

    def geometry(self):
        """(geo [K][14], hmin [K], codes [K]) as the kernels read them (built on the device)."""
        geo, hmin = np.zeros((self.K, 14)), np.zeros(self.K)
        codes = np.zeros(self.K, dtype=np.uint32)
        _check(load().dgx_get_geometry(self._h, _dp(geo), _dp(hmin), codes.ctypes.data_as(C.POINTER(C.c_uint32))))
        return geo, hmin, codes

    def project(self, Uq):
        Uq = np.ascontiguousarray(Uq, dtype=np.float64)
        _check(load().dgx_project(self._h, _dp(Uq)))

    def eval_volume(self):
        Uq = np.zeros((self.K, self.Nq, 4))
        _check(load().dgx_eval_volume(self._h, _dp(Uq)))
        return Uq

    def l2_error(self, exact, primitive=False):
        """Relative L2 error per primitive field (rho, u, v, p) of the current state at the volume
        points against exact values there, [K][Nq][4] conserved (or primitive)."""
        ex = np.ascontiguousarray(exact, dtype=np.float64).reshape(self.K, self.Nq, 4)
        err = np.zeros(4)
        _check(load().dgx_l2_error(self._h, _dp(ex), 1 if primitive else 0, _dp(err)))
        return err

    def write_vtk(self, filename, u=None, gamma=1.4, resolution_factor=3):
        """Legacy-VTK file `filename`.vtk of the state u ([K][Np][4], default: the current one);
        VTKWriter::write_euler_solution (S/utils/vtk_writer.cpp:72-176)."""
        ua = None if u is None else np.ascontiguousarray(u, dtype=np.float64).reshape(self.K, self.Np, 4)
        _check(load().dgx_write_vtk(self._h, None if ua is None else _dp(ua), str(filename).encode(), float(gamma),
                                    int(resolution_factor)))

    # hot path
    def residual(self, u=None):
        r = np.zeros((self.K, self.Np, 4))
        if u is None:
            _check(load().dgx_residual(self._h, None, _dp(r)))
        else:
            u = np.ascontiguousarray(u, dtype=np.float64)
            _check(load().dgx_residual(self._h, _dp(u), _dp(r)))
        return r

    def apply_minv(self, r):
        r = np.array(r, dtype=np.float64, order="C")
        _check(load().dgx_apply_minv(self._h, _dp(r)))
        return r

    def rk3_steps(self, dt, nsteps):
        _check(load().dgx_rk3_steps(self._h, dt, nsteps))

    def rk3_run(self, dt, nsteps):
        """Up to nsteps steps with the reference's per-step NaN abort: returns (steps taken,
        nan), the context holding the state of the last step taken."""
        done = C.c_int32(0)
        st = load().dgx_rk3_run(self._h, dt, nsteps, C.byref(done))
        if st == DGX_E_NONFINITE:
            return done.value, True
        _check(st)
        return done.value, False

    def rk3_steps_async(self, dt, nsteps):
        _check(load().dgx_rk3_steps_async(self._h, dt, nsteps))

    def synchronize(self):
        _check(load().dgx_synchronize(self._h))

    def rk3_profile(self, dt, nsteps):
        s, t = C.c_double(), C.c_double()
        _check(load().dgx_rk3_profile(self._h, dt, nsteps, C.byref(s), C.byref(t)))
        return s.value, t.value

    # diagnostics
    def conservation_sums(self):
        o = np.zeros(4)
        _check(load().dgx_conservation_sums(self._h, _dp(o)))
        return o

    def max_cfl(self, dt):
        o = C.c_double()
        _check(load().dgx_max_cfl(self._h, dt, C.byref(o)))
        return o.value

    def density_range(self):
        o = np.zeros(2)
        _check(load().dgx_density_range(self._h, _dp(o)))
        return o

    def kinetic_energy(self):
        o = np.zeros(1)
        _check(load().dgx_kinetic_energy(self._h, _dp(o)))
        return float(o[0])

    # discrete adjoint (Euler)
    def residual_vjp(self, lam, u=None):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        out = np.zeros((self.K, self.Np, 4))
        up = None if u is None else _dp(np.ascontiguousarray(u, dtype=np.float64))
        _check(load().dgx_residual_vjp(self._h, up, _dp(lam), _dp(out)))
        return out

    def rhs_vjp(self, lam, u=None):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        out = np.zeros((self.K, self.Np, 4))
        up = None if u is None else _dp(np.ascontiguousarray(u, dtype=np.float64))
        _check(load().dgx_rhs_vjp(self._h, up, _dp(lam), _dp(out)))
        return out

    def kinetic_energy_grad(self):
        J = np.zeros(1)
        g = np.zeros((self.K, self.Np, 4))
        _check(load().dgx_kinetic_energy_grad(self._h, _dp(J), _dp(g)))
        return float(J[0]), g

    def rk3_vjp(self, dt, nsteps, ckpt_every, lambda_T=None):
        """lambda_0 = (du_T/du_0)^T lambda_T from the current state; leaves u_T in the context.
        lambda_T None seeds with the kinetic-energy gradient at u_T."""
        out = np.zeros((self.K, self.Np, 4))
        lt = None if lambda_T is None else _dp(np.ascontiguousarray(lambda_T, dtype=np.float64))
        _check(load().dgx_rk3_vjp(self._h, dt, nsteps, ckpt_every, lt, _dp(out)))
        return out

    def rk3_vjp_profile(self, dt, nsteps, ckpt_every):
        """(forward ms, checkpointed forward sweep ms, reverse sweep ms) of an nsteps run."""
        f, a, r = C.c_double(), C.c_double(), C.c_double()
        _check(load().dgx_rk3_vjp_profile(self._h, dt, nsteps, ckpt_every, C.byref(f), C.byref(a), C.byref(r)))
        return f.value, a.value, r.value


class _Borrowed(Context):
    """A Context view of one batch instance (the batch owns the device buffers)."""

    def __init__(self, handle, order, batch):
        self._h = handle
        self._batch = batch  # keep the owner alive
        self.order = order
        info = Info()
        _check(load().dgx_info(handle, C.byref(info)))
        self.info = info
        self.K, self.Np, self.Nq, self.Nfq = info.n_elements, info.n_modes, info.n_volume_points, info.n_face_points
        self.precision = ["fp64", "fp32"][info.precision]

    def close(self):
        self._h = None


class Batch:
    """Many independent Euler instances stepped together (one persistent launch per RK stage,
    graph-captured).  meshes: list of (verts, tris); orders: list of p."""

    def __init__(self, meshes, orders, gamma=1.4, precision="fp32", basis="dubiner", periodic=3):
        L = load()
        n = len(meshes)
        self._keep = []
        descs = (MeshDesc * n)()
        spaces = (SpaceDesc * n)()
        for i, ((v, t), p) in enumerate(zip(meshes, orders)):
            v = np.ascontiguousarray(v, dtype=np.float64)
            t = np.ascontiguousarray(t, dtype=np.int32)
            nb = connect(v, t, periodic)
            bt = np.full(4, -1, dtype=np.int32)
            bv = np.zeros((4, 4))
            self._keep += [v, t, nb, bt, bv]
            descs[i] = MeshDesc(len(v), _dp(v), len(t), _ip(t), _ip(nb), 4, _ip(bt), _dp(bv))
            spaces[i] = SpaceDesc(p, BASIS[basis], 0)
        phys = Physics(0, gamma, 0.0, 0.72, 5.0, 0)
        h = C.c_void_p()
        _check(L.dgx_batch_create(n, descs, spaces, C.byref(phys), PRECISION[precision], C.byref(h)))
        self._h = h
        self.instances = []
        for i in range(n):
            ch = C.c_void_p()
            _check(L.dgx_batch_instance(h, i, C.byref(ch)))
            self.instances.append(_Borrowed(ch, orders[i], self))

    def close(self):
        if getattr(self, "_h", None):
            for c in self.instances:
                c.close()
            load().dgx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_dt(self, dts):
        d = np.ascontiguousarray(dts, dtype=np.float64)
        _check(load().dgx_batch_set_dt(self._h, _dp(d)))

    def rk3_steps(self, nsteps):
        _check(load().dgx_batch_rk3_steps(self._h, nsteps))

    def profile(self, nsteps, sequential=False):
        ms = C.c_double()
        _check(load().dgx_batch_profile(self._h, nsteps, int(sequential), C.byref(ms)))
        return ms.value

    def tuning(self):
        t = BatchTuning()
        _check(load().dgx_batch_get_tuning(self._h, C.byref(t)))
        return t

    def set_tuning(self, layout="concurrent", chunk=(4, 4, 4, 4), blocks_per_cu=4):
        t = BatchTuning({"fused": 0, "per_order": 1, "concurrent": 2}[layout], (C.c_int32 * 4)(*chunk), blocks_per_cu)
        _check(load().dgx_batch_set_tuning(self._h, C.byref(t)))

    def autotune(self, trial_steps=4):
        t = BatchTuning()
        _check(load().dgx_batch_autotune(self._h, trial_steps, C.byref(t)))
        return t


def stable_dt(ctx, cfl=0.3):
    """dt = CFL h_min / ((2p+1) max(|u|+c)) for the current state: the reference's CFL
    measure (compute_max_cfl) scaled by the DG stability factor 1/(2p+1)."""
    lam = ctx.max_cfl(1.0)
    if not (0.0 < lam < float("inf")):
        raise ValueError("the state has no finite positive wave speed (set the state first)")
    return cfl / ((2 * ctx.order + 1) * lam)


# ----------------------------------------------------------------------------------------------
# reference-API mirror
# ----------------------------------------------------------------------------------------------
class DGAssembler:
    """DGAssembler(mesh, weak_form).assemble_euler_residual(u) -> M dU/dt."""

    def __init__(self, ctx):
        self.ctx = ctx

    def assemble_euler_residual(self, u_coeffs):
        return self.ctx.residual(u_coeffs)


class CompressibleDGSolver:
    """EulerDGSolver / NavierStokesDGSolver: solve(ic, T_final, dt, save_every) returns the
    flattened frames (K, 4 Np) with column i*4+v (flatten_frames,
    S/solver/compressible_solver_base.cpp:340-368) under the reference's frame and step-count
    rules (see solve)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.assembler = DGAssembler(ctx)

    def project_initial_condition(self, ic):
        xy = self.ctx.volume_points()
        self.ctx.project(ic(xy))
        return self.ctx.get_state()

    def apply_mass_inv(self, r):
        return self.ctx.apply_minv(r)

    def time_step_ssp_rk3(self, u_n, dt):
        self.ctx.set_state(u_n)
        self.ctx.rk3_steps(dt, 1)
        return self.ctx.get_state()

    def compute_max_cfl(self, u, dt):
        self.ctx.set_state(u)
        return self.ctx.max_cfl(dt)

    def compute_density_range(self, u):
        self.ctx.set_state(u)
        return tuple(self.ctx.density_range())

    def solve(self, ic, T_final, dt, save_every):
        """run_time_integration (S/solver/compressible_solver_base.cpp:201-338): the step count of
        its `while (t < T_final) { ...; t += dt; }` loop, a frame at step 0 and at the top of
        every iteration whose step is a positive multiple of save_every (so only when another
        step follows), the NaN abort after each step, and a last frame when the final step
        count is not a multiple of save_every."""
        if save_every <= 0:
            raise ValueError("save_every must be positive")
        n_steps = reference_step_count(T_final, dt)
        self.project_initial_condition(ic)
        frames = [self.ctx.get_state()]
        step = 0
        while step < n_steps:
            if step % save_every == 0 and step > 0:
                frames.append(self.ctx.get_state())
            chunk = min(save_every - step % save_every, n_steps - step)
            done, nan = self.ctx.rk3_run(dt, chunk)
            step += done
            if nan:
                break
        if step % save_every != 0:
            frames.append(self.ctx.get_state())
        self.steps_taken = step
        return [f.reshape(self.ctx.K, self.ctx.Np * 4) for f in frames]


def max_elements(order, vertices_per_element=3):
    """Largest element count a context of this order and element type accepts."""
    n = C.c_int32(0)
    _check(load().dgx_max_elements(order, vertices_per_element, C.byref(n)))
    return n.value


def reference_step_count(T_final, dt):
    """Iterations of the reference's `while (t < T_final) { ...; t += dt; }` with t starting at
    0.0 (compressible_solver_base.cpp:242,258,284): FP64 accumulation of t can add a step beyond
    ceil(T_final / dt), e.g. 11 steps for T_final = 1, dt = 0.1."""
    if not (dt > 0.0) or not math.isfinite(T_final):
        raise ValueError("dt must be positive and T_final finite")
    t, n = 0.0, 0
    while t < T_final:
        t += dt
        n += 1
    return n


EulerDGSolver = CompressibleDGSolver
NavierStokesDGSolver = CompressibleDGSolver
