// Device helpers shared by the matrix-core kernels (dgx_stage_mfma.hip, dgx_vjp_mfma.hip):
// the 16x16x4 FP32 product, buffer-resource loads and the Euler ghost state.
#pragma once
#include "dgx_mfma.hpp"
#include "dgx_stage_impl.hpp"

namespace dgx {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma16x16x4(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// bfloat16 product with FP32 accumulation, K = 16: lane i + 16 g supplies A[i][4g .. 4g+3] and
// B[4g .. 4g+3][i] as two 32-bit words of packed pairs (low half first).
typedef short v4s __attribute__((ext_vector_type(4)));
typedef unsigned int v2u __attribute__((ext_vector_type(2)));
__device__ __forceinline__ v4f mfma16x16x16bf16(unsigned a0, unsigned a1, unsigned b0, unsigned b1, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(v4s, v2u{a0, a1}),
                                                      __builtin_bit_cast(v4s, v2u{b0, b1}), c, 0, 0, 0);
}
// two floats rounded to bfloat16 (nearest even, v_cvt_pk_bf16_f32) and packed
__device__ __forceinline__ unsigned pack_bf16x2(float lo, float hi) {
    return unsigned(__builtin_bit_cast(unsigned short, (__bf16)lo)) |
           (unsigned(__builtin_bit_cast(unsigned short, (__bf16)hi)) << 16);
}

// Ghost state of a boundary face point from its boundary-data row (bc_row: type, then the
// Euler ghost data) (boundary_face_residual, euler_weak_formulation.cpp:208-233); a face
// without a condition (type < 0) mirrors the interior state.
__device__ __forceinline__ void ghost_state(const float* Uo, float nx, float ny, const double* bcr, float gam,
                                            float* Un) {
    const int type = int(bcr[0]);
    if (type < 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) Un[v] = Uo[v];
    } else if (type == DGX_BC_FAR_FIELD) {
#pragma unroll
        for (int v = 0; v < 4; ++v) Un[v] = float(bcr[1 + v]);
    } else if (type == DGX_BC_SLIP_WALL) {
        const float ri = rcp(Uo[0]);
        const float w1 = Uo[1] * ri, w2 = Uo[2] * ri;
        const float W[4] = {Uo[0], w1, w2, (gam - 1.0f) * (Uo[3] - 0.5f * Uo[0] * (w1 * w1 + w2 * w2))};
        const float vn = W[1] * nx + W[2] * ny;
        const float Wg[4] = {W[0], W[1] - 2.0f * vn * nx, W[2] - 2.0f * vn * ny, W[3]};
        prim2cons(Wg, gam, Un);
    } else {
        const float Wb[4] = {float(bcr[1]), float(bcr[2]), float(bcr[3]), float(bcr[4])};
        prim2cons(Wb, gam, Un);
    }
}

// Twice the Rusanov flux, 2 H(A, B, n) = F(A).n + F(B).n - s (B - A): the same arithmetic as
// rusanov() with the factor 1/2 left to the caller's face scale.
__device__ __forceinline__ void rusanov2(const float* A, const float* B, float nx, float ny, float g, float* H) {
    const float ria = rcp(A[0]), rib = rcp(B[0]);
    const float ua = A[1] * ria, va = A[2] * ria, ub = B[1] * rib, vb = B[2] * rib;
    const float pa = (g - 1.0f) * (A[3] - 0.5f * A[0] * (ua * ua + va * va));
    const float pb = (g - 1.0f) * (B[3] - 0.5f * B[0] * (ub * ub + vb * vb));
    const float una = ua * nx + va * ny, unb = ub * nx + vb * ny;
    const float ca = dsqrt(g * pa * ria), cb = dsqrt(g * pb * rib);
    const float s = ref_max(fabsf(una) + ca, fabsf(unb) + cb);
    const float fa0 = A[0] * una, fb0 = B[0] * unb;
    const float fa1 = A[1] * una + pa * nx, fb1 = B[1] * unb + pb * nx;
    const float fa2 = A[2] * una + pa * ny, fb2 = B[2] * unb + pb * ny;
    const float fa3 = (A[3] + pa) * una, fb3 = (B[3] + pb) * unb;
    H[0] = (fa0 + fb0) - s * (B[0] - A[0]);
    H[1] = (fa1 + fb1) - s * (B[1] - A[1]);
    H[2] = (fa2 + fb2) - s * (B[2] - A[2]);
    H[3] = (fa3 + fb3) - s * (B[3] - A[3]);
}

// Trace record of a face point for the Rusanov flux, formed once by the element that owns the
// point: the normal flux F(U).n along the owner's outward normal and the wave speed |u.n| + c,
// with the arithmetic of rusanov2 for one side.  The normals of the two sides of a face are
// exact negatives, so a neighbour's record gives F(U_nb).n = -Fn_nb.
__device__ __forceinline__ void trace_record(const float* U, float nx, float ny, float g, float* Fn, float& s) {
    const float ri = rcp(U[0]);
    const float u = U[1] * ri, v = U[2] * ri;
    const float p = (g - 1.0f) * (U[3] - 0.5f * U[0] * (u * u + v * v));
    const float un = u * nx + v * ny;
    s = fabsf(un) + dsqrt(g * p * ri);
    Fn[0] = U[0] * un;
    Fn[1] = U[1] * un + p * nx;
    Fn[2] = U[2] * un + p * ny;
    Fn[3] = (U[3] + p) * un;
}

// Flux of one face point in the canonical (L) orientation, scaled for this element's lift;
// bcr is the point's boundary-data row on a boundary face.
__device__ __forceinline__ void face_point_flux(const float* Uo, float* Un, uint32_t c, const double* bcr, float nx,
                                                float ny, float sf, float gam, float* Hs) {
    const bool boundary = (c & 3u) == kFaceBoundary;
    const bool isL = (c >> 3) & 1u;
    if (boundary) ghost_state(Uo, nx, ny, bcr, gam, Un);
    // a boundary face without a boundary condition contributes nothing
    // (S/solver/assembler.cpp:233-243)
    const float sfe = (boundary && bcr[0] < 0.0) ? 0.0f : sf;
    const bool own_is_a = isL || boundary;
    const float sgn = own_is_a ? 1.0f : -1.0f;
    float Aq[4], Bq[4], H[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        Aq[v] = own_is_a ? Uo[v] : Un[v];
        Bq[v] = own_is_a ? Un[v] : Uo[v];
    }
    rusanov(Aq, Bq, sgn * nx, sgn * ny, gam, H);
#pragma unroll
    for (int v = 0; v < 4; ++v) Hs[v] = sgn * sfe * H[v];
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned bytes) {
    // raw buffer (stride 0), range-checked against `bytes`; dword 3 as for every gfx9 target
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, int(bytes), 0x00020000);
}
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
__device__ __forceinline__ v4u load_x4(__amdgpu_buffer_rsrc_t r, unsigned voff) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0);
}
__device__ __forceinline__ double load_f64(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

}  // namespace
}  // namespace dgx
