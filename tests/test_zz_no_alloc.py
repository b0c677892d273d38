"""dgx_conservation_sums_device (the entry point that allocates nothing: partial sums in a fixed
device array, result written to caller-provided device memory) against dgx_conservation_sums
on the same contexts, two calls in a row, the second on a smaller mesh, for triangles in FP64
and FP32 and for quadrilaterals."""
import ctypes as C

import numpy as np
import pytest

import dgx
import refcpu as R

pytestmark = pytest.mark.gpu


def _fn():
    return dgx.load().dgx_conservation_sums_device


def _ctx(kind, n, precision):
    if kind == "tri":
        verts, els = dgx.split_quad_mesh(n, 10.0)
        ctx = dgx.Context(verts, els, 3, basis="dubiner", precision=precision)
    else:
        verts, els = dgx.rect_quad_mesh(0.0, 10.0, 0.0, 10.0, 10.0 / n)
        ctx = dgx.Context(verts, els, 2, precision=precision)
    ctx.project(R.vortex_state(ctx.volume_points()))
    return ctx


def _hip():
    """The HIP runtime libdgx.so itself uses (a second runtime in the process, such as the one
    torch bundles, would not see the library's device state)."""
    dgx.load()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64.so" in line}
    assert len(paths) == 1, paths
    h = C.CDLL(paths.pop())
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def _device_sums(ctx, out, hip):
    st = _fn()(ctx._h, out)
    assert st == dgx.DGX_OK, dgx.load().dgx_last_error()
    ctx.synchronize()
    got = np.full(4, np.nan)
    assert hip.hipMemcpy(got.ctypes.data, out, 32, 2) == 0  # device to host
    return got


@pytest.mark.parametrize("kind,precision", [("tri", "fp64"), ("tri", "fp32"), ("quad", "fp64")])
def test_device_sums_match_the_host_entry_point(kind, precision):
    hip = _hip()
    out = C.c_void_p()
    assert hip.hipMalloc(C.byref(out), 32) == 0  # the caller's output buffer
    try:
        for n in (48, 16):  # the second call on a smaller mesh reuses the same device scratch
            ctx = _ctx(kind, n, precision)
            ctx.rk3_steps(dgx.stable_dt(ctx, 0.3), 3)
            got = _device_sums(ctx, out, hip)
            want = ctx.conservation_sums()
            # the same partial sums added in the same order: equal bit for bit
            assert np.array_equal(got, want), (n, got, want)
            assert np.all(np.isfinite(got)) and abs(got[0]) > 1.0
            ctx.close()
    finally:
        hip.hipFree(out)


def test_rejects_null_arguments():
    assert _fn()(None, None) == dgx.DGX_E_INVALID
