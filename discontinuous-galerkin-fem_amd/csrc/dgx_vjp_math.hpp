// Pointwise pieces of the Euler stage adjoint shared by the per-lane and the matrix-core VJP
// kernels: transposed directional-flux Jacobian and transposed Rusanov Jacobian.
#pragma once
#include "dgx_internal.hpp"

namespace dgx {
namespace {

template <typename Real>
__device__ __forceinline__ Real rcp_(Real x) { return Real(1) / x; }
template <>
__device__ __forceinline__ float rcp_<float>(float x) { return __builtin_amdgcn_rcpf(x); }
// a / b and sqrt(x): correctly rounded in FP64; in FP32 the hardware reciprocal and square root
// (1 ulp, v_rcp_f32 / v_sqrt_f32), as the forward FP32 kernels evaluate them
template <typename Real>
__device__ __forceinline__ Real div_(Real a, Real b) { return a / b; }
template <>
__device__ __forceinline__ float div_<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <typename Real>
__device__ __forceinline__ Real sqrt_(Real x) { return sqrt(x); }
template <>
__device__ __forceinline__ float sqrt_<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

// Ubar += (dFn/dU)^T psi + sbar * d(|un| + c)/dU, Fn = (r un, mx un + p nx, my un + p ny,
// (E+p) un), un = (u nx + v ny) for the (not necessarily unit) direction n.
template <typename Real>
__device__ __forceinline__ void flux_dir_adjoint(const Real* U, Real nx, Real ny, const Real* psi,
                                                 Real sbar, Real g, Real* Ubar) {
    const Real gm1 = g - Real(1);
    Real r = U[0], mx = U[1], my = U[2], E = U[3];
    Real ri = rcp_(r);
    Real u = mx * ri, v = my * ri;
    Real k = Real(0.5) * (u * u + v * v);
    Real p = gm1 * (E - r * k);
    Real un = u * nx + v * ny;
    Real unbar = psi[0] * r + psi[1] * mx + psi[2] * my + psi[3] * (E + p);
    Real pbar = psi[1] * nx + psi[2] * ny + psi[3] * un;
    Real rbar = psi[0] * un, mxbar = psi[1] * un, mybar = psi[2] * un, Ebar = psi[3] * un;
    if (sbar != Real(0)) {
        // s = |un| + c, c = sqrt(g p / r)
        Real c = sqrt_(g * p * ri);
        unbar += sbar * (un >= Real(0) ? Real(1) : Real(-1));
        pbar += div_(sbar * g, Real(2) * c * r);
        rbar -= div_(sbar * c, Real(2) * r);
    }
    Real ubar = unbar * nx, vbar = unbar * ny;
    Ebar += gm1 * pbar;
    rbar -= gm1 * k * pbar;
    ubar -= gm1 * r * u * pbar;
    vbar -= gm1 * r * v * pbar;
    mxbar += ubar * ri;
    mybar += vbar * ri;
    rbar -= (ubar * u + vbar * v) * ri;
    Ubar[0] += rbar;
    Ubar[1] += mxbar;
    Ubar[2] += mybar;
    Ubar[3] += Ebar;
}

template <typename Real>
__device__ __forceinline__ Real wave_speed(const Real* U, Real nx, Real ny, Real g) {
    Real ri = rcp_(U[0]);
    Real u = U[1] * ri, v = U[2] * ri;
    Real p = (g - Real(1)) * (U[3] - Real(0.5) * U[0] * (u * u + v * v));
    return fabs(u * nx + v * ny) + sqrt_(g * p * ri);
}

// Ubar_own = (dH/dU_own)^T psi for H(A, B, n) = 1/2 (Fn(A) + Fn(B)) - 1/2 s (B - A),
// s = max(sA, sB) (A branch at ties).  own_is_A selects the side.
// (rusanov_adjoint_s: the two wave speeds given)
template <typename Real>
__device__ __forceinline__ void rusanov_adjoint_s(const Real* A, const Real* B, Real nx, Real ny, const Real* psi,
                                                  bool own_is_A, Real sA, Real sB, Real g, Real* Ubar) {
    bool from_A = sA >= sB;
    Real s = from_A ? sA : sB;
    Real jump = psi[0] * (B[0] - A[0]) + psi[1] * (B[1] - A[1]) + psi[2] * (B[2] - A[2]) + psi[3] * (B[3] - A[3]);
    Real half_psi[4] = {Real(0.5) * psi[0], Real(0.5) * psi[1], Real(0.5) * psi[2], Real(0.5) * psi[3]};
    const Real* own = own_is_A ? A : B;
    Real sbar = (own_is_A == from_A) ? Real(-0.5) * jump : Real(0);
    flux_dir_adjoint(own, nx, ny, half_psi, sbar, g, Ubar);
    Real sgn = own_is_A ? Real(0.5) : Real(-0.5);
#pragma unroll
    for (int v = 0; v < 4; ++v) Ubar[v] += sgn * s * psi[v];
}
template <typename Real>
__device__ __forceinline__ void rusanov_adjoint(const Real* A, const Real* B, Real nx, Real ny,
                                                const Real* psi, bool own_is_A, Real g, Real* Ubar) {
    rusanov_adjoint_s(A, B, nx, ny, psi, own_is_A, wave_speed(A, nx, ny, g), wave_speed(B, nx, ny, g), g, Ubar);
}

}  // namespace
}  // namespace dgx
