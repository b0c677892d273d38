// Private: the context object behind the dgx C ABI, device buffers and error plumbing shared by
// the host translation units (dgx_api.cpp, dgx_batch_api.cpp).
#pragma once
#include <algorithm>
#include <cmath>
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/dgx.h"
#include "dgx_internal.hpp"

using namespace dgx;

namespace dgx_detail {

extern thread_local std::string g_error;

struct Fail : std::runtime_error {
    dgx_status code;
    Fail(dgx_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Fail(DGX_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
inline dgx_status guarded(F&& f) {
    try {
        f();
        return DGX_OK;
    } catch (const Fail& e) {
        g_error = e.what();
        return e.code;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return DGX_E_INVALID;
    } catch (const std::bad_alloc&) {
        g_error = "out of host memory";
        return DGX_E_NOMEM;
    } catch (const std::exception& e) {
        g_error = e.what();
        return DGX_E_INVALID;
    }
}

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t bytes) {
        release();
        if (bytes == 0) return;
        hip_check(hipMalloc(&p, bytes), "hipMalloc");
        n = bytes;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

}  // namespace dgx_detail

using namespace dgx_detail;

struct dgx_ctx {
    int K = 0, Kp = 0, p = 1, rule = 0, np = 3, nq = 3, nfq = 2, basis = 1, precision = 0, physics = 0;
    int nf = 3;  // faces (= vertices) per element: 3 triangles, 4 quadrilaterals
    bool qtensor = false;  // quads: the four-lanes-per-element Euler stage kernel
    double gamma = 1.4, mu = 0, prandtl = 0.72, sigma0 = 5;
    uint32_t quirks = 0;
    int device = 0;
    int n_bfaces = 0;
    // boundary faces (element, local face, condition row; n_bcs = no condition) in the order of
    // their running index, the conditions' types, constant values and optional functions
    struct BFace { int e, f, bci; };
    std::vector<BFace> h_bfaces;
    std::vector<int> h_bctype;
    std::vector<double> h_bcval;
    std::vector<std::pair<dgx_bc_fn, void*>> h_bcfn;
    bool has_slip = false;  // a boundary face uses a slip-wall condition
    double dt = 0.0;        // time step of this instance inside a batch
    double phi0 = 1.0;      // value of the constant basis function
    Off4 uref{}, off{};     // FP32 storage offset: point values uref, constant-mode coefficient off
    static constexpr Off4 kNoOff{};
    size_t rsz = 8;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // side_stream(), created on first use
    DevBuf u, u1, u2, geo, nbr, code, hmin, bc, tabL, tabR, tab64, minv, pm, qw, stage, partial;
    DevBuf tabML, tabMR;  // MFMA stage kernel lane tables (RK/RHS and residual variants)
    DevBuf tabVL, tabVR;  // MFMA stage adjoint lane tables (same two variants)
    DevBuf tabNT;         // MFMA Navier-Stokes face-trace and external-face tables
    DevBuf minve;         // quads: per-element inverse mass matrices [Np*Np][Kp]
    DevBuf ub;            // the state at the start of a checked run (dgx_rk3_run), allocated on first use
    // matrix-core Euler tile kernel: FP32 geometry (RK/RHS and M dU/dt variants), face
    // connections, external-face lists and the largest per-tile count of interior external faces
    DevBuf geo32L, geo32R, conn, ext, ext_off, ext_nint, halo;
    int halo_cap = 0;
    std::vector<double> h_minv, h_qw, h_verts;  // h_verts [K][nf][2] in the context's element order
    // The context numbers the elements in its own order (triangles: Morton order of the
    // centroids, so that a 64-element tile is a compact patch of the mesh); every per-element
    // host array crosses the C ABI in the caller's order.  perm[e] = caller index of element e,
    // iperm its inverse; both empty when the orders agree.
    std::vector<int32_t> perm, iperm;
    int dev_of(int ec) const { return iperm.empty() ? ec : iperm[ec]; }
    // caller-order rows of width w -> context order (h itself when the orders agree)
    template <typename T>
    const T* rows_in(const T* h, size_t w, std::vector<T>& tmp) const {
        if (perm.empty() || !h) return h;
        tmp.resize((size_t)K * w);
        for (int e = 0; e < K; ++e) std::copy(h + (size_t)perm[e] * w, h + (size_t)perm[e] * w + w, tmp.data() + (size_t)e * w);
        return tmp.data();
    }
    // context-order rows of width w, in place -> caller order
    template <typename T>
    void rows_out(T* h, size_t w) const {
        if (perm.empty() || !h) return;
        std::vector<T> tmp(h, h + (size_t)K * w);
        for (int e = 0; e < K; ++e) std::copy(tmp.data() + (size_t)e * w, tmp.data() + (size_t)e * w + w, h + (size_t)perm[e] * w);
    }
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    ~dgx_ctx() {
        for (DevBuf* b : {&u, &u1, &u2, &geo, &nbr, &code, &hmin, &bc, &tabL, &tabR, &tab64, &minv, &pm, &qw, &stage,
                          &partial, &tabML, &tabMR, &tabVL, &tabVR, &tabNT, &minve, &ub, &geo32L, &geo32R, &conn,
                          &ext, &ext_off, &ext_nint, &halo})
            b->release();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream2) (void)hipStreamDestroy(stream2);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Choose the storage offset from a representative state U (FP32 only): rounded to FP32 so
    // that the kernels add back exactly what pack subtracted (phi0 is a power of two).
    void set_offset(const double U[4]) {
        if (precision == DGX_FP64) return;
        for (int v = 0; v < 4; ++v) {
            uref.v[v] = std::isfinite(U[v]) ? double(float(U[v])) : 0.0;
            off.v[v] = uref.v[v] / phi0;
        }
    }

    size_t state_bytes() const { return (size_t)np * 4 * Kp * rsz; }
    int red_blocks() const { return (K + 255) / 256; }

    void upload(DevBuf& b, const void* src, size_t bytes) {
        b.alloc(bytes);
        hip_check(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    template <typename T>
    void upload_real(DevBuf& b, const std::vector<double>& v) {
        std::vector<T> t(v.begin(), v.end());
        upload(b, t.data(), t.size() * sizeof(T));
    }

    // A second stream of default priority (the adjoint's recomputation beside its reverse
    // sweep).  Streams share the process's few hardware queues once those are all in use, and
    // two streams on one queue run one after the other; a stream with a CU mask gets a queue of
    // its own, so the side stream is made with a mask that enables every CU.
    hipStream_t side_stream() {
        if (!stream2) {
            hipDeviceProp_t prop;
            hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
            std::vector<uint32_t> mask((std::max(prop.multiProcessorCount, 1) + 31) / 32, 0xffffffffu);
            hip_check(hipExtStreamCreateWithCUMask(&stream2, (uint32_t)mask.size(), mask.data()), "hipStreamCreate");
        }
        return stream2;
    }

    void stage_launch(const void* uk, const void* un, void* out, int mode, double dt, double ca, double cb,
                      hipStream_t on = nullptr) {
        const hipStream_t stream = on ? on : this->stream;
        StageParams sp{};
        sp.uk = uk;
        sp.un = un;
        sp.out = out;
        sp.geo = geo.as<double>();
        sp.nbr = nbr.as<int32_t>();
        sp.code = code.as<uint32_t>();
        sp.tab = (mode == MODE_RESIDUAL) ? tabR.p : tabL.p;
        sp.tabm = (mode == MODE_RESIDUAL) ? tabMR.p : tabML.p;
        sp.tabt = tabNT.p;
        sp.bc = bc.as<double>();
        sp.minve = minve.as<double>();
        sp.tab64 = tab64.as<double>();
        sp.K = K;
        sp.Kp = Kp;
        sp.dt = dt;
        sp.ca = ca;
        sp.cb = cb;
        sp.gamma = gamma;
        sp.mu = mu;
        sp.prandtl = prandtl;
        sp.sigma0 = sigma0;
        sp.mode = mode;
        sp.ns_face_quirk = (quirks & DGX_QUIRK_NS_FACE_GRAD_REFERENCE) ? 1 : 0;
        sp.ns_avg_ref = (quirks & DGX_QUIRK_NS_VISCOUS_AVERAGE_REFERENCE) ? 1 : 0;
        sp.kappa = ((quirks & DGX_QUIRK_NS_HEAT_FLUX_REFERENCE) ? 1.0 : -1.0) * mu * gamma / (prandtl * (gamma - 1.0));
        for (int v = 0; v < 4; ++v) {
            sp.uref[v] = uref.v[v];
            sp.uoff[v] = off.v[v];
        }
        set_tile_params(sp, mode);
        if (nf == 4 && qtensor)
            hip_check(launch_stage_quad_tensor(precision, p, sp, stream), "stage kernel launch");
        else if (nf == 4)
            hip_check(launch_stage_quad(precision, physics, p, sp.ns_face_quirk != 0, sp, stream), "stage kernel launch");
        else
            hip_check(launch_stage(precision, physics, p, rule, sp, stream), "stage kernel launch");
    }

    // the matrix-core tile kernels' arrays (null when the context has none)
    void set_tile_params(StageParams& sp, int mode) const {
        sp.f.gamma = float(sp.gamma);
        sp.f.mu = float(sp.mu);
        sp.f.lam = (-2.0f * sp.f.mu) / 3.0f;  // as viscous_lambda() forms it on the device
        sp.f.kappa = float(sp.kappa);
        sp.f.sig0 = float(sp.sigma0 * sp.mu * (p + 1) * (p + 1));
        sp.f.dt = float(sp.dt);
        sp.f.ca = float(sp.ca);
        sp.f.cb = float(sp.cb);
        for (int v = 0; v < 4; ++v) {
            sp.f.uref[v] = float(sp.uref[v]);
            sp.f.uoff[v] = float(sp.uoff[v]);
        }
        sp.geo32 = (mode == MODE_RESIDUAL ? geo32R : geo32L).as<float>();
        sp.conn = conn.as<uint32_t>();
        sp.ext = ext.as<int32_t>();
        sp.ext_off = ext_off.as<int32_t>();
        sp.ext_nint = ext_nint.as<int32_t>();
        sp.halo = halo.as<int32_t>();
        sp.halo_cap = halo_cap;
    }

    // the matrix-core stage adjoint's tile arrays and FP32 scalars (null arrays without them)
    void set_vjp_tile_params(VjpParams& vp, int mode) const {
        vp.geo32 = (mode == MODE_RESIDUAL ? geo32R : geo32L).as<float>();
        vp.conn = conn.as<uint32_t>();
        vp.ext = ext.as<int32_t>();
        vp.ext_off = ext_off.as<int32_t>();
        vp.ext_nint = ext_nint.as<int32_t>();
        vp.halo = halo.as<int32_t>();
        vp.halo_cap = halo_cap;
        vp.f.gamma = float(vp.gamma);
        vp.f.dt = float(vp.dt);
        vp.f.ca = float(vp.ca);
        vp.f.cb = float(vp.cb);
        for (int v = 0; v < 4; ++v) vp.f.uref[v] = float(vp.uref[v]);
    }

    // One SSP-RK3 step (SSP_RK::step_rk3, I/solver/time_stepping.hpp:39-55):
    //   u1 = u + dt L(u); u2 = 3/4 u + 1/4 (u1 + dt L(u1)); u = 1/3 u + 2/3 (u2 + dt L(u2)).
    void rk3_step(double dt) {
        stage_launch(u.p, u.p, u1.p, MODE_RK, dt, 0.0, 1.0);
        stage_launch(u1.p, u.p, u2.p, MODE_RK, dt, 0.75, 0.25);
        stage_launch(u2.p, u.p, u1.p, MODE_RK, dt, 1.0 / 3.0, 2.0 / 3.0);
        std::swap(u, u1);
    }

    // NaN entries of the state's coefficients (the reference's per-element hasNaN()); infinite
    // entries are not counted, as there
    double nan_count() {
        std::vector<double> h = reduce(3);
        double n = 0.0;
        for (int b = 0; b < red_blocks(); ++b) n += h[b * 4];
        return n;
    }

    std::vector<double> reduce(int kind) {
        int nb = red_blocks();
        if (nf == 4)
            hip_check(launch_reduce_quad(kind, precision, np, nq, u.p, tab64.as<double>(), geo.as<double>(),
                                         hmin.as<double>(), K, Kp, gamma, uref, partial.as<double>(), nb, stream),
                      "reduce kernel launch");
        else
            hip_check(launch_reduce(kind, precision, p, rule, u.p, tab64.p, geo.as<double>(), hmin.as<double>(),
                                    qw.as<double>(), K, Kp, gamma, uref, off, partial.as<double>(), nb, stream),
                      "reduce kernel launch");
        std::vector<double> h((size_t)nb * 4);
        hip_check(hipMemcpyAsync(h.data(), partial.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream), "copy");
        hip_check(hipStreamSynchronize(stream), "sync");
        return h;
    }
};
