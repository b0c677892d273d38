// Fused DG stage kernel for gfx950: volume residual + face residuals (Rusanov, and for
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
// Layout: state u[(i*4+v)*Kp + e] (mode-major, element fastest) so a wave's 64 elements read
// 256 contiguous bytes per coefficient; FP64 geometry geo[row*Kp + e]; tables are wave-uniform
// (scalar loads) except the neighbour-face tables, staged in LDS because the neighbour's
// local face differs between lanes.
//
// Element-centric face handling: each element evaluates all three of its faces, recomputing
// the neighbour's trace from the neighbour's coefficients.  The flux of a face is always formed
// in the orientation of its canonical (L) element, with the L-side normal, and enters the R
// element with the opposite sign, so both copies come from the same arithmetic.
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
__device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_sqrtf(x); }

// Block -> element-tile remap so that consecutive tiles (mesh neighbours) land on the same
// XCD and share its L2 (blocks are dealt round-robin over the 8 XCDs).
__device__ __forceinline__ int xcd_tile(int bid, int nblk) {
    if ((nblk & 7) != 0) return bid;
    return (bid & 7) * (nblk >> 3) + (bid >> 3);
}

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
    Real s = fmax(fabs(una) + ca, fabs(unb) + cb);
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
// navier_stokes_weak_formulation.cpp:36-98), gas constant 1.  g[v*2+d] = d U_v / d x_d.
template <typename Real>
__device__ __forceinline__ void viscous_flux(const Real* U, const Real* g, Real gam, Real mu,
                                             Real kappa, Real* Fv, Real* Gv) {
    Real rho = U[0];
    Real ri_safe = rcp(fmax(rho, Real(1e-12)));
    Real ri = rcp(rho);
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
    Real lam = Real(-2) * mu / Real(3);
    Real txx = Real(2) * mu * ux + lam * div;
    Real tyy = Real(2) * mu * vy + lam * div;
    Real txy = mu * (uy + vx);
    Real qx = -kappa * Tx, qy = -kappa * Ty;
    Fv[0] = Real(0); Fv[1] = txx; Fv[2] = txy; Fv[3] = u * txx + v * txy + qx;
    Gv[0] = Real(0); Gv[1] = txy; Gv[2] = tyy; Gv[3] = u * txy + v * tyy + qy;
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
// With the reference's face-gradient read (quirk) every slot is populated; otherwise only
// d == c, and the zero slots are skipped at compile time.
template <int NP, bool QUIRK, typename Real, typename TabFn>
__device__ __forceinline__ void face_gradient(const Real (&u)[NP][4], const double* A, TabFn tab,
                                              Real* g) {
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
                    for (int v = 0; v < 4; ++v) s[v] = fma(t, u[i][v], s[v]);
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

template <int P, int RULE, typename Real, bool NS, bool QUIRK>
__global__ __launch_bounds__(kBlock) void stage_kernel(StageParams sp) {
    constexpr int NP = np_of(P);
    constexpr int NQ = nq_of(P, RULE);
    constexpr int NFQ = P + 1;
    // table offsets (TabLayout, compile-time)
    constexpr int O_phiV = 0;
    constexpr int O_PX = O_phiV + NQ * NP;
    constexpr int O_PY = O_PX + NQ * NP;
    constexpr int O_phiF = O_PY + NQ * NP;
    constexpr int O_LF = O_phiF + 3 * NFQ * NP;
    constexpr int O_DXV = O_LF + 3 * NFQ * NP;
    constexpr int O_DYV = O_DXV + NQ * NP;
    constexpr int O_FG = O_DYV + NQ * NP;  // face-gradient tables [3][NFQ][8][NP]
    constexpr int NFG = NS ? 3 * NFQ * 8 * NP : 1;

    const Real* __restrict__ tab = static_cast<const Real*>(sp.tab);
    const Real* __restrict__ uk = static_cast<const Real*>(sp.uk);
    const Real* __restrict__ un = static_cast<const Real*>(sp.un);
    Real* __restrict__ out = static_cast<Real*>(sp.out);
    const double* __restrict__ geo = sp.geo;
    const int Kp = sp.Kp;

    // neighbour-side tables in LDS (lane-varying local face / point order)
    __shared__ Real s_phiF[3 * NFQ * NP];
    __shared__ Real s_FG[NFG];
    for (int t = threadIdx.x; t < 3 * NFQ * NP; t += kBlock) s_phiF[t] = tab[O_phiF + t];
    if constexpr (NS)
        for (int t = threadIdx.x; t < NFG; t += kBlock) s_FG[t] = tab[O_FG + t];
    __syncthreads();

    const int e = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (e >= sp.K) return;

    const Real gam = Real(sp.gamma);
    const Real mu = Real(sp.mu);
    const Real kappa = Real(sp.mu * sp.gamma / (sp.prandtl * (sp.gamma - 1.0)));
    const bool residual_mode = sp.mode == MODE_RESIDUAL;

    Real u[NP][4];
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) u[i][v] = uk[(i * 4 + v) * Kp + e];

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
        Real U[4] = {0, 0, 0, 0};
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
#pragma unroll
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
            for (int v = 0; v < 4; ++v) Uo[q][v] = Real(0);
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                Real ph = tab[O_phiF + (f * NFQ + q) * NP + i];
#pragma unroll
                for (int v = 0; v < 4; ++v) Uo[q][v] = fma(ph, u[i][v], Uo[q][v]);
            }
        }

        Real Un[NFQ][4];
        Real un_c[NP][4];
        double An[4] = {0, 0, 0, 0};
        bool boundary = (fn == int(kFaceBoundary));
        if (!boundary) {
#pragma unroll
            for (int i = 0; i < NP; ++i)
#pragma unroll
                for (int v = 0; v < 4; ++v) un_c[i][v] = uk[(i * 4 + v) * Kp + nb];
#pragma unroll
            for (int q = 0; q < NFQ; ++q) {
                const int qq = flip ? (NFQ - 1 - q) : q;
                const Real* row = &s_phiF[(fn * NFQ + qq) * NP];
#pragma unroll
                for (int v = 0; v < 4; ++v) Un[q][v] = Real(0);
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    Real ph = row[i];
#pragma unroll
                    for (int v = 0; v < 4; ++v) Un[q][v] = fma(ph, un_c[i][v], Un[q][v]);
                }
            }
            if constexpr (NS) {
#pragma unroll
                for (int k = 0; k < 4; ++k) An[k] = geo[(G_A00 + k) * Kp + nb];
            }
        } else {
            // ghost state from the boundary condition (boundary_face_residual, :208-233)
            const double* bcr = sp.bc + 5 * nb;
            const int type = int(bcr[0]);
#pragma unroll
            for (int q = 0; q < NFQ; ++q) {
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
        const bool no_bc = boundary && sp.bc[5 * nb] < 0.0;
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
#pragma unroll
            for (int q = 0; q < NFQ; ++q) {
                Real go[8], gn[8];
                auto own_tab = [&](int slot, int i) {
                    return tab[O_FG + ((f * NFQ + q) * 8 + slot) * NP + i];
                };
                face_gradient<NP, QUIRK>(u, A, own_tab, go);
                Real Fvo[4], Gvo[4], Fvn[4] = {0, 0, 0, 0}, Gvn[4] = {0, 0, 0, 0};
                viscous_flux(Uo[q], go, gam, mu, kappa, Fvo, Gvo);
                Real Fn[4];
                if (!boundary) {
                    const int qq = flip ? (NFQ - 1 - q) : q;
                    auto nb_tab = [&](int slot, int i) { return s_FG[((fn * NFQ + qq) * 8 + slot) * NP + i]; };
                    face_gradient<NP, QUIRK>(un_c, An, nb_tab, gn);
                    viscous_flux(Un[q], gn, gam, mu, kappa, Fvn, Gvn);
                    // Fn = 1/2 Fv_L.n - 1/2 Fv_R.n - sigma (U_R - U_L), n = L normal (:192-196)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        Real fo = Fvo[v] * cnx + Gvo[v] * cny;
                        Real fnb = Fvn[v] * cnx + Gvn[v] * cny;
                        Real fL = isL ? fo : fnb, fR = isL ? fnb : fo;
                        Real UL = isL ? Uo[q][v] : Un[q][v], UR = isL ? Un[q][v] : Uo[q][v];
                        Fn[v] = Real(0.5) * fL - Real(0.5) * fR - sigma * (UR - UL);
                    }
                } else {
                    // boundary: Fn = Fv_L.n - sigma (U_bc - U_L), U_bc from the BC data taken as
                    // conserved except for NO_SLIP (:249-264)
                    const double* bcr = sp.bc + 5 * nb;
                    const int type = int(bcr[0]);
                    Real Ub[4];
                    if (type == DGX_BC_NO_SLIP_WALL) {
                        Real Wb[4] = {Real(bcr[1]), Real(bcr[2]), Real(bcr[3]), Real(bcr[4])};
                        prim2cons(Wb, gam, Ub);
                    } else {
#pragma unroll
                        for (int v = 0; v < 4; ++v) Ub[v] = Real(bcr[1 + v]);
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        Fn[v] = (Fvo[v] * nx + Gvo[v] * ny) - sigma * (Ub[v] - Uo[q][v]);
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
        const Real dt = Real(sp.dt), ca = Real(sp.ca), cb = Real(sp.cb);
        if (sp.ca == 0.0) {
#pragma unroll
            for (int i = 0; i < NP; ++i)
#pragma unroll
                for (int v = 0; v < 4; ++v) out[(i * 4 + v) * Kp + e] = u[i][v] + dt * acc[i][v];
        } else {
#pragma unroll
            for (int i = 0; i < NP; ++i)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    Real unv = un[(i * 4 + v) * Kp + e];
                    out[(i * 4 + v) * Kp + e] = ca * unv + cb * (u[i][v] + dt * acc[i][v]);
                }
        }
    } else {
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) out[(i * 4 + v) * Kp + e] = acc[i][v];
    }
}

template <int P, int RULE, typename Real>
hipError_t launch_p(int physics, bool quirk, const StageParams& sp, hipStream_t s) {
    dim3 grid((sp.K + kBlock - 1) / kBlock), block(kBlock);
    if (physics == DGX_EULER)
        hipLaunchKernelGGL((stage_kernel<P, RULE, Real, false, false>), grid, block, 0, s, sp);
    else if (quirk)
        hipLaunchKernelGGL((stage_kernel<P, RULE, Real, true, true>), grid, block, 0, s, sp);
    else
        hipLaunchKernelGGL((stage_kernel<P, RULE, Real, true, false>), grid, block, 0, s, sp);
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_real(int physics, int p, int rule, const StageParams& sp, hipStream_t s) {
    bool q = sp.ns_face_quirk != 0;
    if (rule == 0) {
        switch (p) {
        case 1: return launch_p<1, 0, Real>(physics, q, sp, s);
        case 2: return launch_p<2, 0, Real>(physics, q, sp, s);
        case 3: return launch_p<3, 0, Real>(physics, q, sp, s);
        case 4: return launch_p<4, 0, Real>(physics, q, sp, s);
        }
    } else {
        switch (p) {
        case 1: return launch_p<1, 1, Real>(physics, q, sp, s);
        case 2: return launch_p<2, 1, Real>(physics, q, sp, s);
        case 3: return launch_p<3, 1, Real>(physics, q, sp, s);
        case 4: return launch_p<4, 1, Real>(physics, q, sp, s);
        }
    }
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_stage(int precision, int physics, int p, int rule, const StageParams& sp,
                        hipStream_t s) {
    if (sp.K <= 0) return hipSuccess;
    if (precision == DGX_FP64) return launch_real<double>(physics, p, rule, sp, s);
    return launch_real<float>(physics, p, rule, sp, s);
}

}  // namespace dgx
