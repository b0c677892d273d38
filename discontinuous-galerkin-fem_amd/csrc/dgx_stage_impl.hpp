// Fused DG stage kernel for gfx950 (implementation header, included once per
// precision/physics family by dgx_stage_*.hip): volume residual + face residuals (Rusanov, and for
// Navier-Stokes the viscous volume/face terms) + inverse mass + SSP-RK3 stage update, one
// thread per element, one launch per RK stage.
//
// Reference behaviour (paths relative to the reference root):
//   volume      S/weak_forms/euler_weak_formulation.cpp:72-115
//   faces       S/weak_forms/euler_weak_formulation.cpp:117-250
//   viscous     S/weak_forms/navier_stokes_weak_formulation.cpp:36-278
//   assembly    S/solver/assembler.cpp:137-279
//   M^-1, RK3   S/solver/compressible_solver_base.cpp:54-61,106-112, I/solver/time_stepping.hpp:39-55
//
// Layout: state u[(i*Kp + e)*4 + v] (mode-major rows, the 4 variables of an element contiguous),
// so a wave's 64 elements read 1 KiB contiguous per mode; FP64 geometry geo[row*Kp + e]; tables
// are wave-uniform
// (scalar loads) except the neighbour-face tables, staged in LDS because the neighbour's
// local face differs between lanes.
//
// Element-centric face handling: each element evaluates all three of its faces, recomputing
// the neighbour's trace from the neighbour's coefficients.  The flux of a face is always formed
// in the orientation of its canonical (L) element, with the L-side normal, and enters the R
// element with the opposite sign, so both copies come from the same arithmetic.
#pragma once
#include "dgx_internal.hpp"

namespace dgx {
namespace {

template <typename Real>
__device__ __forceinline__ Real rcp(Real x) { return Real(1) / x; }
template <>
__device__ __forceinline__ float rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

template <typename Real>
__device__ __forceinline__ Real dsqrt(Real x) { return sqrt(x); }
template <>
__device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }  // v_sqrt_f32, 1 ulp

// std::max(a, b) as the reference evaluates it, (a < b) ? b : a: a NaN first operand is
// returned, a NaN second operand is dropped (fmax drops either), so a state that goes NaN does
// so at the same step as in the reference.
template <typename Real>
__device__ __forceinline__ Real ref_max(Real a, Real b) { return (a < b) ? b : a; }


// Rusanov flux H(A, B, n) (rusanov_flux, euler_weak_formulation.cpp:51-70), with the normal
// fluxes written as F.n = (rho un, rho u un + p nx, rho v un + p ny, (E+p) un).
template <typename Real>
__device__ __forceinline__ void rusanov(const Real* A, const Real* B, Real nx, Real ny, Real g,
                                        Real* H) {
    const Real half = Real(0.5);
    Real ria = rcp(A[0]), rib = rcp(B[0]);
    Real ua = A[1] * ria, va = A[2] * ria, ub = B[1] * rib, vb = B[2] * rib;
    Real pa = (g - Real(1)) * (A[3] - half * A[0] * (ua * ua + va * va));
    Real pb = (g - Real(1)) * (B[3] - half * B[0] * (ub * ub + vb * vb));
    Real una = ua * nx + va * ny, unb = ub * nx + vb * ny;
    Real ca = dsqrt(g * pa * ria), cb = dsqrt(g * pb * rib);
    Real s = ref_max(fabs(una) + ca, fabs(unb) + cb);
    Real fa0 = A[0] * una, fb0 = B[0] * unb;
    Real fa1 = A[1] * una + pa * nx, fb1 = B[1] * unb + pb * nx;
    Real fa2 = A[2] * una + pa * ny, fb2 = B[2] * unb + pb * ny;
    Real fa3 = (A[3] + pa) * una, fb3 = (B[3] + pb) * unb;
    H[0] = half * (fa0 + fb0) - half * s * (B[0] - A[0]);
    H[1] = half * (fa1 + fb1) - half * s * (B[1] - A[1]);
    H[2] = half * (fa2 + fb2) - half * s * (B[2] - A[2]);
    H[3] = half * (fa3 + fb3) - half * s * (B[3] - A[3]);
}

// Viscous fluxes (compute_primitive_gradients + compute_viscous_fluxes,
// navier_stokes_weak_formulation.cpp:36-98), gas constant 1.  g[v*2+d] = d U_v / d x_d;
// lam = -2 mu / 3 (viscous_lambda(mu), formed once by the caller).
template <typename Real>
__device__ __forceinline__ Real viscous_lambda(Real mu) { return Real(-2) * mu / Real(3); }
template <typename Real>
__device__ __forceinline__ void viscous_flux_l(const Real* U, const Real* g, Real gam, Real mu, Real lam,
                                               Real kappa, Real* Fv, Real* Gv) {
    Real rho = U[0];
    Real ri = rcp(rho);
    // 1 / max(rho, 1e-12) with the reference's max (a NaN density passes through)
    Real ri_safe = rho < Real(1e-12) ? rcp(Real(1e-12)) : ri;
    Real u = U[1] * ri, v = U[2] * ri;
    Real p = (gam - Real(1)) * (U[3] - Real(0.5) * rho * (u * u + v * v));
    Real ux = (g[2] - u * g[0]) * ri_safe, uy = (g[3] - u * g[1]) * ri_safe;
    Real vx = (g[4] - v * g[0]) * ri_safe, vy = (g[5] - v * g[1]) * ri_safe;
    Real ksq = u * u + v * v;
    Real px = (gam - Real(1)) * (g[6] - Real(0.5) * ksq * g[0] - rho * (u * ux + v * vx));
    Real py = (gam - Real(1)) * (g[7] - Real(0.5) * ksq * g[1] - rho * (u * uy + v * vy));
    Real rsq_i = rcp(rho * rho);
    Real Tx = (rho * px - p * g[0]) * rsq_i, Ty = (rho * py - p * g[1]) * rsq_i;
    Real div = ux + vy;
    Real txx = Real(2) * mu * ux + lam * div;
    Real tyy = Real(2) * mu * vy + lam * div;
    Real txy = mu * (uy + vx);
    Real qx = -kappa * Tx, qy = -kappa * Ty;
    Fv[0] = Real(0); Fv[1] = txx; Fv[2] = txy; Fv[3] = u * txx + v * txy + qx;
    Gv[0] = Real(0); Gv[1] = txy; Gv[2] = tyy; Gv[3] = u * txy + v * tyy + qy;
}
template <typename Real>
__device__ __forceinline__ void viscous_flux(const Real* U, const Real* g, Real gam, Real mu,
                                             Real kappa, Real* Fv, Real* Gv) {
    viscous_flux_l(U, g, gam, mu, viscous_lambda(mu), kappa, Fv, Gv);
}

template <typename Real>
__device__ __forceinline__ void prim2cons(const Real* W, Real g, Real* U) {
    U[0] = W[0];
    U[1] = W[0] * W[1];
    U[2] = W[0] * W[2];
    U[3] = W[3] / (g - Real(1)) + Real(0.5) * W[0] * (W[1] * W[1] + W[2] * W[2]);
}

// Gradient of U at one face point from the 8-slot face-gradient table
// T[(c*2 + d)*2 + k][i]:  g[v*2+c] = sum_{d,k} A[d][k] sum_i T[c,d,k][i] u[i][v].
// With the reference's face-gradient read (quirk) every slot is populated -- the constant
// mode's row included, so it takes the full constant-mode coefficient (stored value plus the
// storage offset `off`); otherwise only d == c, and the zero slots are skipped at compile time.
template <int NP, bool QUIRK, typename Real, typename TabFn>
__device__ __forceinline__ void face_gradient(const Real (&u)[NP][4], const double* A, TabFn tab,
                                              const Real (&off)[4], Real* g) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        Real acc[4] = {0, 0, 0, 0};
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            if (!QUIRK && d != c) continue;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                Real s[4] = {0, 0, 0, 0};
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    Real t = tab(((c * 2 + d) * 2 + k), i);
#pragma unroll
                    for (int v = 0; v < 4; ++v) s[v] = fma(t, (QUIRK && i == 0) ? u[i][v] + off[v] : u[i][v], s[v]);
                }
                Real a = Real(A[d * 2 + k]);
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[v] = fma(a, s[v], acc[v]);
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) g[v * 2 + c] = acc[v];
    }
}

// Compile-time table offsets (TabLayout).
template <int P, int RULE, bool NS>
struct StageTab {
    static constexpr int NP = np_of(P);
    static constexpr int NQ = nq_of(P, RULE);
    static constexpr int NFQ = P + 1;
    static constexpr int phiV = 0;
    static constexpr int PX = phiV + NQ * NP;
    static constexpr int PY = PX + NQ * NP;
    static constexpr int phiF = PY + NQ * NP;
    static constexpr int LF = phiF + 3 * NFQ * NP;
    static constexpr int DXV = LF + 3 * NFQ * NP;
    static constexpr int DYV = DXV + NQ * NP;
    static constexpr int FG = DYV + NQ * NP;  // face-gradient tables [3][NFQ][8][NP]
    static constexpr int NFG = NS ? 3 * NFQ * 8 * NP : 1;
    static constexpr int NF = 3 * NFQ * NP;
};

// Neighbour-side tables into LDS (the neighbour's local face and point order vary per lane):
// s_phiF [face][point][mode], s_phiFT [face][mode][point], s_FG face gradients (NS).
template <int P, int RULE, typename Real, bool NS>
__device__ __forceinline__ void stage_tables_to_lds(const Real* __restrict__ tab, Real* s_phiF, Real* s_phiFT,
                                                    Real* s_FG, int tid, int nthreads) {
    using T = StageTab<P, RULE, NS>;
    for (int t = tid; t < T::NF; t += nthreads) {
        s_phiF[t] = tab[T::phiF + t];
        const int f = t / (T::NFQ * T::NP), q = (t / T::NP) % T::NFQ, i = t % T::NP;
        s_phiFT[(f * T::NP + i) * T::NFQ + q] = tab[T::phiF + t];
    }
    if constexpr (NS)
        for (int t = tid; t < T::NFG; t += nthreads) s_FG[t] = tab[T::FG + t];
}

// One element's fused stage: volume + faces + M^-1 + RK update (or RHS / residual output).
template <int P, int RULE, typename Real, bool NS, bool QUIRK>
__device__ __forceinline__ void stage_element(const StageParams& sp, const int e, const double dt_e,
                                              const Real* s_phiF, const Real* s_phiFT, const Real* s_FG) {
    using T = StageTab<P, RULE, NS>;
    constexpr int NP = T::NP;
    constexpr int NQ = T::NQ;
    constexpr int NFQ = T::NFQ;
    constexpr int O_phiV = T::phiV;
    constexpr int O_PX = T::PX;
    constexpr int O_PY = T::PY;
    constexpr int O_phiF = T::phiF;
    constexpr int O_LF = T::LF;
    constexpr int O_DXV = T::DXV;
    constexpr int O_DYV = T::DYV;
    constexpr int O_FG = T::FG;

    const Real* __restrict__ tab = static_cast<const Real*>(sp.tab);
    const Real* __restrict__ uk = static_cast<const Real*>(sp.uk);
    const Real* __restrict__ un = static_cast<const Real*>(sp.un);
    Real* __restrict__ out = static_cast<Real*>(sp.out);
    const double* __restrict__ geo = sp.geo;
    const int Kp = sp.Kp;

    const Real gam = Real(sp.gamma);
    const Real mu = Real(sp.mu);
    const Real kappa = Real(sp.kappa);
    const bool residual_mode = sp.mode == MODE_RESIDUAL;
    // point values start from the storage offset (zero in FP64)
    const Real ur[4] = {Real(sp.uref[0]), Real(sp.uref[1]), Real(sp.uref[2]), Real(sp.uref[3])};
    const Real uo[4] = {Real(sp.uoff[0]), Real(sp.uoff[1]), Real(sp.uoff[2]), Real(sp.uoff[3])};

    Real u[NP][4];
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) u[i][v] = uk[sidx(i, v, e, Kp)];

    double A[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) A[k] = geo[(G_A00 + k) * Kp + e];
    const double det = geo[G_DET * Kp + e];
    const Real a00 = Real(A[0]), a01 = Real(A[1]), a10 = Real(A[2]), a11 = Real(A[3]);
    // volume terms carry |J| only for M dU/dt; with M^-1 = M_ref^-1/|J| folded they cancel
    const Real sv = residual_mode ? Real(det) : Real(1);
    const double rdet = 1.0 / det;

    Real acc[NP][4];
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[i][v] = Real(0);

    // ---------------- volume: U_q, flux, contravariant projection -------------------------
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        Real U[4] = {ur[0], ur[1], ur[2], ur[3]};
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            Real ph = tab[O_phiV + q * NP + i];
#pragma unroll
            for (int v = 0; v < 4; ++v) U[v] = fma(ph, u[i][v], U[v]);
        }
        Real ri = rcp(U[0]);
        Real uu = U[1] * ri, vv = U[2] * ri;
        Real p = (gam - Real(1)) * (U[3] - Real(0.5) * U[0] * (uu * uu + vv * vv));
        // contravariant velocities along the rows of dxi_dx^T
        Real c1 = uu * a00 + vv * a10, c2 = uu * a01 + vv * a11;
        Real Fh[4], Gh[4];
        Fh[0] = U[0] * c1;               Gh[0] = U[0] * c2;
        Fh[1] = U[1] * c1 + p * a00;     Gh[1] = U[1] * c2 + p * a01;
        Fh[2] = U[2] * c1 + p * a10;     Gh[2] = U[2] * c2 + p * a11;
        Fh[3] = (U[3] + p) * c1;         Gh[3] = (U[3] + p) * c2;
        if constexpr (NS) {
            Real gx[4] = {0, 0, 0, 0}, gy[4] = {0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                Real dx = tab[O_DXV + q * NP + i], dy = tab[O_DYV + q * NP + i];
#pragma unroll
                for (int v = 0; v < 4; ++v) { gx[v] = fma(dx, u[i][v], gx[v]); gy[v] = fma(dy, u[i][v], gy[v]); }
            }
            Real g[8];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                g[v * 2] = a00 * gx[v] + a01 * gy[v];
                g[v * 2 + 1] = a10 * gx[v] + a11 * gy[v];
            }
            Real Fv[4], Gv[4];
            viscous_flux(U, g, gam, mu, kappa, Fv, Gv);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                Fh[v] -= a00 * Fv[v] + a10 * Gv[v];
                Gh[v] -= a01 * Fv[v] + a11 * Gv[v];
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) { Fh[v] *= sv; Gh[v] *= sv; }
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            Real px = tab[O_PX + q * NP + i], py = tab[O_PY + q * NP + i];
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[i][v] = fma(px, Fh[v], fma(py, Gh[v], acc[i][v]));
        }
    }

    // ---------------- faces ----------------------------------------------------------------
    const uint32_t code = sp.code[e];
    // the face index stays a (wave-uniform) runtime value: table reads remain scalar loads and
    // only one face's traces are live at a time
#pragma unroll 1
    for (int f = 0; f < 3; ++f) {
        const uint32_t c = (code >> (4 * f)) & 15u;
        const int fn = int(c & 3u);
        const bool flip = (c >> 2) & 1u;
        const bool isL = (c >> 3) & 1u;
        const int nb = sp.nbr[f * Kp + e];
        const double len = geo[(G_L0 + f) * Kp + e];
        const Real nx = Real(geo[(G_N0X + 2 * f) * Kp + e]);
        const Real ny = Real(geo[(G_N0X + 2 * f + 1) * Kp + e]);
        const Real sf = residual_mode ? Real(len) : Real(len * rdet);

        // own trace
        Real Uo[NFQ][4];
#pragma unroll
        for (int q = 0; q < NFQ; ++q) {
#pragma unroll
            for (int v = 0; v < 4; ++v) Uo[q][v] = ur[v];
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                Real ph = tab[O_phiF + (f * NFQ + q) * NP + i];
#pragma unroll
                for (int v = 0; v < 4; ++v) Uo[q][v] = fma(ph, u[i][v], Uo[q][v]);
            }
        }

        Real Un[NFQ][4];
        Real un_c[NS ? NP : 1][4];
        double An[4] = {0, 0, 0, 0};
        bool boundary = (fn == int(kFaceBoundary));
        if (!boundary) {
            if constexpr (!NS) {
                // stream the neighbour's coefficients mode by mode: only its trace stays live
#pragma unroll
                for (int q = 0; q < NFQ; ++q)
#pragma unroll
                    for (int v = 0; v < 4; ++v) Un[q][v] = ur[v];
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    Real cf[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) cf[v] = uk[sidx(i, v, nb, Kp)];
                    const Real* col = &s_phiFT[(fn * NP + i) * NFQ];
                    Real ph[NFQ];
#pragma unroll
                    for (int q = 0; q < NFQ; ++q) ph[q] = col[q];
#pragma unroll
                    for (int q = 0; q < NFQ; ++q) {
                        const Real t = flip ? ph[NFQ - 1 - q] : ph[q];
#pragma unroll
                        for (int v = 0; v < 4; ++v) Un[q][v] = fma(t, cf[v], Un[q][v]);
                    }
                }
            } else {
#pragma unroll
                for (int i = 0; i < NP; ++i)
#pragma unroll
                    for (int v = 0; v < 4; ++v) un_c[i][v] = uk[sidx(i, v, nb, Kp)];
#pragma unroll
                for (int q = 0; q < NFQ; ++q) {
                    const int qq = flip ? (NFQ - 1 - q) : q;
                    const Real* row = &s_phiF[(fn * NFQ + qq) * NP];
#pragma unroll
                    for (int v = 0; v < 4; ++v) Un[q][v] = ur[v];
#pragma unroll
                    for (int i = 0; i < NP; ++i) {
                        Real ph = row[i];
#pragma unroll
                        for (int v = 0; v < 4; ++v) Un[q][v] = fma(ph, un_c[i][v], Un[q][v]);
                    }
                }
            }
            if constexpr (NS) {
#pragma unroll
                for (int k = 0; k < 4; ++k) An[k] = geo[(G_A00 + k) * Kp + nb];
            }
        } else {
            // ghost state from the boundary condition at each face point (boundary_face_residual,
            // :208-233)
            const int type = int(bc_row(sp.bc, nb, NFQ, 0)[0]);
#pragma unroll
            for (int q = 0; q < NFQ; ++q) {
                const double* bcr = bc_row(sp.bc, nb, NFQ, q);
                Real W[4];
                Real ri = rcp(Uo[q][0]);
                W[0] = Uo[q][0]; W[1] = Uo[q][1] * ri; W[2] = Uo[q][2] * ri;
                W[3] = (gam - Real(1)) * (Uo[q][3] - Real(0.5) * W[0] * (W[1] * W[1] + W[2] * W[2]));
                if (type < 0) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) Un[q][v] = Uo[q][v];
                } else if (type == DGX_BC_FAR_FIELD) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) Un[q][v] = Real(bcr[1 + v]);
                } else if (type == DGX_BC_SLIP_WALL) {
                    Real vn = W[1] * nx + W[2] * ny;
                    Real Wg[4] = {W[0], W[1] - Real(2) * vn * nx, W[2] - Real(2) * vn * ny, W[3]};
                    prim2cons(Wg, gam, Un[q]);
                } else {
                    Real Wb[4] = {Real(bcr[1]), Real(bcr[2]), Real(bcr[3]), Real(bcr[4])};
                    prim2cons(Wb, gam, Un[q]);
                }
            }
        }

        // a boundary face without a boundary condition contributes nothing (the assembler
        // skips faces whose bc_euler is unset, S/solver/assembler.cpp:233-243)
        const bool no_bc = boundary && bc_row(sp.bc, nb, NFQ, 0)[0] < 0.0;
        const Real sfe = no_bc ? Real(0) : sf;
        // canonical orientation: A-side = L element, normal = L-side outward normal
        const Real sgn = (isL || boundary) ? Real(1) : Real(-1);
        const Real cnx = sgn * nx, cny = sgn * ny;

        Real Hs[NFQ][4];
#pragma unroll
        for (int q = 0; q < NFQ; ++q) {
            Real Aq[4], Bq[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                Aq[v] = (isL || boundary) ? Uo[q][v] : Un[q][v];
                Bq[v] = (isL || boundary) ? Un[q][v] : Uo[q][v];
            }
            Real H[4];
            rusanov(Aq, Bq, cnx, cny, gam, H);
#pragma unroll
            for (int v = 0; v < 4; ++v) Hs[q][v] = sgn * sfe * H[v];
        }

        if constexpr (NS) {
            // viscous face flux; penalty sigma0 mu (p+1)^2 / max(h, 1e-12) (:275-278)
            const Real sigma = Real(sp.sigma0 * sp.mu * (P + 1) * (P + 1) / fmax(len, 1e-12));
            // weights of the two sides' fluxes along the L normal: (1/2, -1/2) the reference's
            // half difference, (-1/2, -1/2) the consistent -{Fv}.n; boundary sign likewise
            const Real wL = sp.ns_avg_ref ? Real(0.5) : Real(-0.5), wR = Real(-0.5);
            const Real wB = sp.ns_avg_ref ? Real(1) : Real(-1);
#pragma unroll
            for (int q = 0; q < NFQ; ++q) {
                Real go[8], gn[8];
                auto own_tab = [&](int slot, int i) {
                    return tab[O_FG + ((f * NFQ + q) * 8 + slot) * NP + i];
                };
                face_gradient<NP, QUIRK>(u, A, own_tab, uo, go);
                Real Fvo[4], Gvo[4], Fvn[4] = {0, 0, 0, 0}, Gvn[4] = {0, 0, 0, 0};
                viscous_flux(Uo[q], go, gam, mu, kappa, Fvo, Gvo);
                Real Fn[4];
                if (!boundary) {
                    const int qq = flip ? (NFQ - 1 - q) : q;
                    auto nb_tab = [&](int slot, int i) { return s_FG[((fn * NFQ + qq) * 8 + slot) * NP + i]; };
                    face_gradient<NP, QUIRK>(un_c, An, nb_tab, uo, gn);
                    viscous_flux(Un[q], gn, gam, mu, kappa, Fvn, Gvn);
                    // Fn = 1/2 Fv_L.n -+ 1/2 Fv_R.n - sigma (U_R - U_L), n = L normal (:192-196)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        Real fo = Fvo[v] * cnx + Gvo[v] * cny;
                        Real fnb = Fvn[v] * cnx + Gvn[v] * cny;
                        Real fL = isL ? fo : fnb, fR = isL ? fnb : fo;
                        Real UL = isL ? Uo[q][v] : Un[q][v], UR = isL ? Un[q][v] : Uo[q][v];
                        Fn[v] = wL * fL + wR * fR - sigma * (UR - UL);
                    }
                } else {
                    // boundary: Fn = Fv_L.n - sigma (U_bc - U_L), U_bc from the BC data taken as
                    // conserved except for NO_SLIP (:249-264)
                    const double* bcr = bc_row(sp.bc, nb, NFQ, q);
                    const int type = int(bcr[0]);
                    Real Ub[4];
                    if (type == DGX_BC_NO_SLIP_WALL) {
                        Real Wb[4] = {Real(bcr[5]), Real(bcr[6]), Real(bcr[7]), Real(bcr[8])};
                        prim2cons(Wb, gam, Ub);
                    } else {
#pragma unroll
                        for (int v = 0; v < 4; ++v) Ub[v] = Real(bcr[5 + v]);
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        Fn[v] = wB * (Fvo[v] * nx + Gvo[v] * ny) - sigma * (Ub[v] - Uo[q][v]);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) Hs[q][v] += sgn * sfe * Fn[v];
            }
        }

        // lift: acc_i += LF[f][q][i] * Hs_q  (LF carries -w_q/2 and M_ref^-1)
#pragma unroll
        for (int q = 0; q < NFQ; ++q)
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                Real l = tab[O_LF + (f * NFQ + q) * NP + i];
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[i][v] = fma(l, Hs[q][v], acc[i][v]);
            }
    }

    // ---------------- output ---------------------------------------------------------------
    if (sp.mode == MODE_RK) {
        const Real dt = Real(dt_e), ca = Real(sp.ca), cb = Real(sp.cb);
        if (sp.ca == 0.0) {
#pragma unroll
            for (int i = 0; i < NP; ++i)
#pragma unroll
                for (int v = 0; v < 4; ++v) out[sidx(i, v, e, Kp)] = u[i][v] + dt * acc[i][v];
        } else {
#pragma unroll
            for (int i = 0; i < NP; ++i)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    Real unv = un[sidx(i, v, e, Kp)];
                    out[sidx(i, v, e, Kp)] = ca * unv + cb * (u[i][v] + dt * acc[i][v]);
                }
        }
    } else {
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) out[sidx(i, v, e, Kp)] = acc[i][v];
    }
}

template <int P, int RULE, typename Real, bool NS, bool QUIRK>
__global__ __launch_bounds__(kBlock) void stage_kernel(StageParams sp) {
    using T = StageTab<P, RULE, NS>;
    __shared__ Real s_phiF[T::NF];
    __shared__ Real s_phiFT[T::NF];
    __shared__ Real s_FG[T::NFG];
    stage_tables_to_lds<P, RULE, Real, NS>(static_cast<const Real*>(sp.tab), s_phiF, s_phiFT, s_FG, threadIdx.x, kBlock);
    __syncthreads();
    const int e = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (e >= sp.K) return;
    stage_element<P, RULE, Real, NS, QUIRK>(sp, e, sp.dt, s_phiF, s_phiFT, s_FG);
}

// Launch of one (precision, physics) family for any p and volume rule.
template <typename Real, bool NS, bool QUIRK>
hipError_t launch_family(int p, int rule, const StageParams& sp, hipStream_t s) {
    dim3 grid((sp.K + kBlock - 1) / kBlock), block(kBlock);
#define DGX_LAUNCH(PP, RR) \
    hipLaunchKernelGGL((stage_kernel<PP, RR, Real, NS, QUIRK>), grid, block, 0, s, sp)
    switch (rule * 8 + p) {
    case 1: DGX_LAUNCH(1, 0); break;
    case 2: DGX_LAUNCH(2, 0); break;
    case 3: DGX_LAUNCH(3, 0); break;
    case 4: DGX_LAUNCH(4, 0); break;
    case 9: DGX_LAUNCH(1, 1); break;
    case 10: DGX_LAUNCH(2, 1); break;
    case 11: DGX_LAUNCH(3, 1); break;
    case 12: DGX_LAUNCH(4, 1); break;
    default: return hipErrorInvalidValue;
    }
#undef DGX_LAUNCH
    return hipGetLastError();
}

}  // namespace
}  // namespace dgx
