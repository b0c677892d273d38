"""CPU checks of the C ABI: the library loads, exports every entry point include/dgx.h
declares, its host-only mesh helpers agree with the oracle, and argument errors come back
as status codes (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import dgx
import refcpu as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "dgx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:dgx_status|void|const char\*)\s+(dgx_\w+)\s*\(", text, flags=re.M)


def test_header_declares_what_the_binding_lists():
    decl = declared_functions()
    assert len(decl) >= 20
    assert set(decl) == set(dgx.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = dgx.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_split_quad_mesh_matches_oracle_generator():
    for n in (2, 5, 16):
        v1, t1 = dgx.split_quad_mesh(n, 10.0)
        v2, t2 = R.split_quad_mesh(n, 10.0)
        assert np.array_equal(t1, t2)
        assert np.array_equal(v1, v2)


@pytest.mark.parametrize("periodic", [0, 1, 2, 3])
def test_connectivity_matches_oracle(periodic):
    v, t = dgx.split_quad_mesh(6, 2.0, -1.0, 0.5)
    nb = dgx.connect(v, t, periodic)
    o = R.Oracle(v, t, 1, periodic=periodic)
    ref = o.neighbors()
    inner = ref[..., 0] >= 0
    assert np.array_equal(nb[inner], ref[inner])
    # boundary faces: (-1, tag) with tags 0..3 = Bottom, Right, Top, Left
    assert (nb[~inner][:, 0] == -1).all()
    assert set(np.unique(nb[~inner][:, 1])) <= {0, 1, 2, 3}
    assert (~inner).sum() == o.n_boundary


def test_connectivity_rejects_bad_meshes():
    v, t = dgx.split_quad_mesh(2)
    bad = t.copy()
    bad[0, 0] = len(v) + 5
    with pytest.raises(dgx.DGXError) as e:
        dgx.connect(v, bad)
    assert e.value.status == dgx.DGX_E_INVALID
    # the same edge used by three triangles
    v3 = np.array([[0, 0], [1, 0], [0, 1], [1, 1], [0, -1.0]])
    t3 = np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]], dtype=np.int32)
    with pytest.raises(dgx.DGXError):
        dgx.connect(v3, t3, 0)


def _create(order=3, basis=1, precision=1, kind=0, mu=0.0):
    v, t = dgx.split_quad_mesh(3)
    nb = dgx.connect(v, t)
    mesh = dgx.MeshDesc(len(v), v.ctypes.data_as(C.POINTER(C.c_double)), len(t),
                        t.ctypes.data_as(C.POINTER(C.c_int32)), nb.ctypes.data_as(C.POINTER(C.c_int32)),
                        0, None, None)
    space = dgx.SpaceDesc(order, basis, 0)
    phys = dgx.Physics(kind, 1.4, mu, 0.72, 5.0, 0)
    h = C.c_void_p()
    st = dgx.load().dgx_create(C.byref(mesh), C.byref(space), C.byref(phys), precision, C.byref(h))
    return st, h, (v, t, nb)


def test_create_validates_arguments_before_touching_a_device():
    assert _create(order=7)[0] == dgx.DGX_E_INVALID
    assert _create(basis=9)[0] == dgx.DGX_E_INVALID
    assert _create(precision=5)[0] == dgx.DGX_E_INVALID
    # the bfloat16 volume variant exists for Euler at p = 3, 4 only
    assert _create(precision=2, order=2)[0] == dgx.DGX_E_UNSUPPORTED
    assert _create(precision=2, kind=1, mu=1e-3)[0] == dgx.DGX_E_UNSUPPORTED
    # NavierStokesWeakFormulation ctor checks
    st, _, _ = _create(kind=1, mu=-1.0)
    assert st == dgx.DGX_E_INVALID
    assert b"viscosity" in dgx.load().dgx_last_error()


def test_no_cpu_fallback_without_a_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    st, h, _ = _create()
    assert st == dgx.DGX_E_DEVICE
    assert not h.value


def test_null_arguments_are_rejected():
    L = dgx.load()
    assert L.dgx_set_state(None, None) == dgx.DGX_E_INVALID
    assert L.dgx_rk3_steps(None, 0.1, 1) == dgx.DGX_E_INVALID
    assert L.dgx_create(None, None, None, 0, None) == dgx.DGX_E_INVALID


@pytest.mark.parametrize("nv", [3, 4])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_max_elements_is_the_32_bit_state_index_boundary(order, nv):
    # state index (i Kp + e) 4 + v < 2^31 for i < Np, e < Kp (K rounded up to 256): the last
    # accepted K keeps it, the next multiple of 256 does not
    np_ = (order + 1) ** 2 if nv == 4 else (order + 1) * (order + 2) // 2
    kmax = dgx.max_elements(order, nv)
    kp = -(-kmax // 256) * 256
    assert kmax % 256 == 0
    assert 4 * np_ * kp <= 2 ** 31
    assert 4 * np_ * (kp + 256) > 2 ** 31
    # quads of order 4 stop near 21.5 M elements, triangles of order 1 near 179 M
    if (order, nv) == (4, 4):
        assert 21_000_000 < kmax < 21_500_000


@pytest.mark.parametrize("nv", [3, 4])
def test_create_rejects_meshes_beyond_the_index_range_first(nv):
    # n_elements one past the limit with tiny arrays: rejected before the arrays are read or
    # anything is allocated (host or device), for triangles and quadrilaterals alike
    order = 4
    kmax = dgx.max_elements(order, nv)
    v = np.zeros((4, 2))
    t = np.zeros((1, nv), dtype=np.int32)
    nb = np.zeros((1, nv, 2), dtype=np.int32)
    mesh = dgx.MeshDesc(len(v), v.ctypes.data_as(C.POINTER(C.c_double)), kmax + 1,
                        t.ctypes.data_as(C.POINTER(C.c_int32)), nb.ctypes.data_as(C.POINTER(C.c_int32)),
                        0, None, None, nv)
    space = dgx.SpaceDesc(order, 2 if nv == 4 else 1, 0)
    phys = dgx.Physics(0, 1.4, 0.0, 0.72, 5.0, 0)
    h = C.c_void_p()
    st = dgx.load().dgx_create(C.byref(mesh), C.byref(space), C.byref(phys), 1, C.byref(h))
    assert st == dgx.DGX_E_UNSUPPORTED and not h.value
    assert b"32 bits" in dgx.load().dgx_last_error()
