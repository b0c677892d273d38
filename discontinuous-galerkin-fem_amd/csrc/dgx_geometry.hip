// Per-element FP64 geometry built on the device: the affine mapping's inverse Jacobian and
// determinant, the outward unit normals and lengths of the three faces (the canonical L side's
// normal and length on both elements of an interior face, negated on the R side), the smallest
// edge length, and the 4-bit face codes (neighbour's local face, reversed point order, L side).
//
// GeometricMapping::compute_mapping (S/reference/mapping.cpp:22-50), DGSpace::compute_face_data
// (S/core/space.cpp:104-137) and compute_face_permutation (:173-228); the element size of
// compute_max_cfl (S/solver/compressible_solver_base.cpp:125-135, Eigen norm()).
//
// The arithmetic is written without contractions so that every entry is the correctly rounded
// result of the same operation sequence on the host: the geometry is bitwise reproducible.
#include "dgx_internal.hpp"

#pragma clang fp contract(off)

namespace dgx {
namespace {

struct Tri {
    double x[3], y[3];
};

__device__ __forceinline__ Tri load_tri(const double* __restrict__ V, const int32_t* __restrict__ T, int e) {
    Tri t;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int id = T[e * 3 + k];
        t.x[k] = V[2 * id];
        t.y[k] = V[2 * id + 1];
    }
    return t;
}

// outward unit normal and length of local face f = edge (v_f, v_{f+1})
__device__ __forceinline__ void face_geom(const Tri& t, int f, double& nx, double& ny, double& len) {
    const int a = f, b = f == 2 ? 0 : f + 1;
    const double tx = t.x[b] - t.x[a], ty = t.y[b] - t.y[a];
    len = sqrt(tx * tx + ty * ty);
    nx = ty / len;
    ny = -tx / len;
    const double cx = (t.x[0] + t.x[1] + t.x[2]) / 3.0, cy = (t.y[0] + t.y[1] + t.y[2]) / 3.0;
    if (nx * (0.5 * (t.x[a] + t.x[b]) - cx) + ny * (0.5 * (t.y[a] + t.y[b]) - cy) < 0) {
        nx = -nx;
        ny = -ny;
    }
}

__device__ __forceinline__ double dist(double ax, double ay, double bx, double by) {
    const double dx = ax - bx, dy = ay - by;
    return sqrt(dx * dx + dy * dy);
}

// caller: the caller's index of each element (null when the context keeps the caller's order);
// the canonical (L) side of an interior face is the element with the smaller caller index, so
// the geometry and codes do not depend on the context's element order.
__global__ __launch_bounds__(256) void geometry_kernel(const double* __restrict__ V, const int32_t* __restrict__ T,
                                                       const int32_t* __restrict__ N, const int32_t* __restrict__ caller,
                                                       int K, int Kp,
                                                       double* __restrict__ geo, double* __restrict__ hmin,
                                                       uint32_t* __restrict__ code, int* __restrict__ err) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Kp) return;
    const size_t sK = (size_t)Kp;
    if (e >= K) {  // padding elements: identity-free zeros, |det J| = 1, h = 1
#pragma unroll
        for (int r = 0; r < kGeoRows; ++r) geo[r * sK + e] = r == G_DET ? 1.0 : 0.0;
        hmin[e] = 1.0;
        code[e] = 0;
        return;
    }
    const Tri t = load_tri(V, T, e);
    const double jt00 = t.x[1] - t.x[0], jt01 = t.y[1] - t.y[0], jt10 = t.x[2] - t.x[0], jt11 = t.y[2] - t.y[0];
    const double det = jt00 * jt11 - jt01 * jt10;
    int bad = fabs(det) < 1e-12 ? 1 : 0;
    const double id = 1.0 / det;
    geo[G_A00 * sK + e] = jt11 * id;
    geo[G_A01 * sK + e] = -jt01 * id;
    geo[G_A10 * sK + e] = -jt10 * id;
    geo[G_A11 * sK + e] = jt00 * id;
    geo[G_DET * sK + e] = fabs(det);
    const double l01 = dist(t.x[1], t.y[1], t.x[0], t.y[0]), l12 = dist(t.x[2], t.y[2], t.x[1], t.y[1]),
                 l20 = dist(t.x[0], t.y[0], t.x[2], t.y[2]);
    hmin[e] = fmin(l01, fmin(l12, l20));

    uint32_t cw = 0;
    for (int f = 0; f < 3; ++f) {
        const int en = N[(e * 3 + f) * 2], fn = N[(e * 3 + f) * 2 + 1];
        double nx, ny, len;
        face_geom(t, f, nx, ny, len);
        if (len < 1e-14) bad |= 2;
        uint32_t cf;
        if (en < 0) {
            cf = kFaceBoundary | (1u << 3);
        } else {
            const int ce = caller ? caller[e] : e, cn = caller ? caller[en] : en;
            const bool isL = (ce < cn) || (ce == cn && f < fn);
            const Tri tn = load_tri(V, T, en);
            const Tri& tL = isL ? t : tn;
            const Tri& tR = isL ? tn : t;
            const int fL = isL ? f : fn, fR = isL ? fn : f;
            double nLx, nLy, lenL, nRx, nRy, lenR;
            face_geom(tL, fL, nLx, nLy, lenL);
            face_geom(tR, fR, nRx, nRy, lenR);
            // point order of the R side's face relative to the L side's
            const int a1 = fL, a2 = fL == 2 ? 0 : fL + 1, b1 = fR, b2 = fR == 2 ? 0 : fR + 1;
            const double tol = 1e-10;
            const bool rev = dist(tL.x[a1], tL.y[a1], tR.x[b2], tR.y[b2]) < tol &&
                             dist(tL.x[a2], tL.y[a2], tR.x[b1], tR.y[b1]) < tol;
            const bool same = dist(tL.x[a1], tL.y[a1], tR.x[b1], tR.y[b1]) < tol &&
                              dist(tL.x[a2], tL.y[a2], tR.x[b2], tR.y[b2]) < tol;
            bool reverse;
            if (rev || same) {
                reverse = rev;
            } else {  // periodic partners: compare the edge directions
                const double t1x = tL.x[a2] - tL.x[a1], t1y = tL.y[a2] - tL.y[a1];
                const double t2x = tR.x[b2] - tR.x[b1], t2y = tR.y[b2] - tR.y[b1];
                const double n1 = sqrt(t1x * t1x + t1y * t1y), n2 = sqrt(t2x * t2x + t2y * t2y);
                reverse = (t1x / n1 * t2x / n2 + t1y / n1 * t2y / n2) < 0.0;
            }
            nx = isL ? nLx : -nLx;
            ny = isL ? nLy : -nLy;
            len = lenL;
            cf = uint32_t(fn) | (reverse ? 4u : 0u) | (isL ? 8u : 0u);
        }
        geo[(G_N0X + 2 * f) * sK + e] = nx;
        geo[(G_N0X + 2 * f + 1) * sK + e] = ny;
        geo[(G_L0 + f) * sK + e] = len;
        cw |= cf << (4 * f);
    }
    code[e] = cw;
    if (bad) atomicOr(err, bad);
}

// The FP32 operands the matrix-core tile kernels consume, formed in FP64 from the FP64 geometry
// exactly as the per-element kernels form them (dxi_dx scaled by sv = |det J| for M dU/dt or 1,
// the lift scale len / |det J| with 1 / |det J| in FP64; dxi_dx and the face lengths as they
// are) and rounded once.
__global__ __launch_bounds__(256) void geometry32_kernel(const double* __restrict__ geo, int K, int Kp, int residual,
                                                         float* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Kp) return;
    const size_t sK = (size_t)Kp;
    if (e >= K) {
        for (int r = 0; r < kGeo32NsRows; ++r) out[r * sK + e] = 0.0f;
        return;
    }
    const double det = geo[G_DET * sK + e];
    const double rdet = 1.0 / det;
    const double sv = residual ? det : 1.0;
    out[G32_A00 * sK + e] = float(sv * geo[G_A00 * sK + e]);
    out[G32_A01 * sK + e] = float(sv * geo[G_A01 * sK + e]);
    out[G32_A10 * sK + e] = float(sv * geo[G_A10 * sK + e]);
    out[G32_A11 * sK + e] = float(sv * geo[G_A11 * sK + e]);
    for (int k = 0; k < 6; ++k) out[(G32_N0X + k) * sK + e] = float(geo[(G_N0X + k) * sK + e]);
    for (int f = 0; f < 3; ++f) {
        const double len = geo[(G_L0 + f) * sK + e];
        out[(G32_S0 + f) * sK + e] = residual ? float(len) : float(len * rdet);
        out[(G32_L0 + f) * sK + e] = float(len);
    }
    for (int k = 0; k < 4; ++k) out[(G32_G00 + k) * sK + e] = float(geo[(G_A00 + k) * sK + e]);
}

}  // namespace

hipError_t launch_geometry32(const double* geo, int K, int Kp, int residual, float* out, hipStream_t s) {
    hipLaunchKernelGGL(geometry32_kernel, dim3((Kp + 255) / 256), dim3(256), 0, s, geo, K, Kp, residual, out);
    return hipGetLastError();
}

hipError_t launch_geometry(const double* V, const int32_t* T, const int32_t* N, const int32_t* caller, int K, int Kp,
                           double* geo, double* hmin, uint32_t* code, int* err, hipStream_t s) {
    const int blocks = (Kp + 255) / 256;
    hipLaunchKernelGGL(geometry_kernel, dim3(blocks), dim3(256), 0, s, V, T, N, caller, K, Kp, geo, hmin, code, err);
    return hipGetLastError();
}

}  // namespace dgx
