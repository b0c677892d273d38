// Discrete adjoint (vector-Jacobian product) of the Euler stage operator for gfx950.
//
// Forward (stage kernel, RK mode):  L_e = sum_q PX_q F^(U_q) + PY_q G^(U_q)
//                                       + sum_f (len_f/|J_e|) sum_q LF_fq s_ef H_f(q)
// with F^, G^ the fluxes along the rows of dxi_dx^T, H_f the Rusanov flux of face f in its
// canonical (L-side) orientation and s_ef = +1 on the L side, -1 on the R side.  This kernel
// computes  ubar_e = (dL/du)^T lambda  restricted to element e's own coefficients, one thread per
// element: the volume term through the transposed directional-flux Jacobian, and each face term
// through the transposed Rusanov Jacobian with respect to e's own trace, weighted by the
// adjoint of the face flux
//   psi(q) = s_e (len/|J_e|) sum_i LF[f][q][i] lambda_e,i + s_n (len/|J_n|) sum_i LF[fn][q'][i] lambda_n,i
// which needs the neighbour's adjoint coefficients and |J_n| (gathered like its state).
// max() in the Rusanov speed takes the A (L-side) branch at ties (a consistent sub-gradient).
// Boundary faces are supported for FAR_FIELD (constant ghost state) and unset conditions.
//
// RK stage adjoint (SSP_RK::step_rk3 transposed, I/solver/time_stepping.hpp:39-55):
//   stage  out = ca un + cb (uk + dt L(uk))
//   =>     lam_uk = cb (lam + dt (dL/du)^T lam) + add,   acc_out = acc_in + ca lam
// Mode RESIDUAL uses the unfolded tables (|J| and len scaling) and returns (dR/du)^T lam.
#include "dgx_mfma.hpp"
#include "dgx_vjp_math.hpp"

namespace dgx {

hipError_t launch_vjp_mfma(int p, const VjpParams& vp, hipStream_t s);

namespace {


template <int P, int RULE, typename Real>
__global__ __launch_bounds__(kBlock) void vjp_kernel(VjpParams sp) {
    constexpr int NP = np_of(P);
    constexpr int NQ = nq_of(P, RULE);
    constexpr int NFQ = P + 1;
    constexpr int O_phiV = 0;
    constexpr int O_PX = O_phiV + NQ * NP;
    constexpr int O_PY = O_PX + NQ * NP;
    constexpr int O_phiF = O_PY + NQ * NP;
    constexpr int O_LF = O_phiF + 3 * NFQ * NP;

    const Real* __restrict__ tab = static_cast<const Real*>(sp.tab);
    const Real* __restrict__ uk = static_cast<const Real*>(sp.uk);
    const Real* __restrict__ lam = static_cast<const Real*>(sp.lam);
    const double* __restrict__ geo = sp.geo;
    const int Kp = sp.Kp;

    __shared__ Real s_phiFT[3 * NP * NFQ];  // [face][mode][point]
    __shared__ Real s_LFT[3 * NP * NFQ];
    for (int t = threadIdx.x; t < 3 * NFQ * NP; t += kBlock) {
        const int f = t / (NFQ * NP), q = (t / NP) % NFQ, i = t % NP;
        s_phiFT[(f * NP + i) * NFQ + q] = tab[O_phiF + t];
        s_LFT[(f * NP + i) * NFQ + q] = tab[O_LF + t];
    }
    __syncthreads();

    const int e = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (e >= sp.K) return;
    const Real g = Real(sp.gamma);
    const bool residual_mode = sp.mode == MODE_RESIDUAL;

    Real u[NP][4], lm[NP][4], ub[NP][4];
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            u[i][v] = uk[sidx(i, v, e, Kp)];
            lm[i][v] = lam[sidx(i, v, e, Kp)];
            ub[i][v] = Real(0);
        }
    const double det = geo[G_DET * Kp + e];
    const Real a00 = Real(geo[G_A00 * Kp + e]), a01 = Real(geo[G_A01 * Kp + e]);
    const Real a10 = Real(geo[G_A10 * Kp + e]), a11 = Real(geo[G_A11 * Kp + e]);
    const Real sv = residual_mode ? Real(det) : Real(1);
    const Real ur[4] = {Real(sp.uref[0]), Real(sp.uref[1]), Real(sp.uref[2]), Real(sp.uref[3])};

    // ---------------- volume ----------------
#pragma unroll 1
    for (int q = 0; q < NQ; ++q) {
        Real U[4] = {ur[0], ur[1], ur[2], ur[3]}, a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            Real ph = tab[O_phiV + q * NP + i];
            Real px = tab[O_PX + q * NP + i] * sv, py = tab[O_PY + q * NP + i] * sv;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                U[v] = fma(ph, u[i][v], U[v]);
                a[v] = fma(px, lm[i][v], a[v]);
                b[v] = fma(py, lm[i][v], b[v]);
            }
        }
        Real Ub[4] = {0, 0, 0, 0};
        flux_dir_adjoint(U, a00, a10, a, Real(0), g, Ub);
        flux_dir_adjoint(U, a01, a11, b, Real(0), g, Ub);
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            Real ph = tab[O_phiV + q * NP + i];
#pragma unroll
            for (int v = 0; v < 4; ++v) ub[i][v] = fma(ph, Ub[v], ub[i][v]);
        }
    }

    // ---------------- faces ----------------
    const uint32_t code = sp.code[e];
#pragma unroll 1
    for (int f = 0; f < 3; ++f) {
        const uint32_t c = (code >> (4 * f)) & 15u;
        const int fn = int(c & 3u);
        const bool flip = (c >> 2) & 1u;
        const bool isL = (c >> 3) & 1u;
        const bool boundary = fn == int(kFaceBoundary);
        const int nb = sp.nbr[f * Kp + e];
        const double len = geo[(G_L0 + f) * Kp + e];
        const Real nx = Real(geo[(G_N0X + 2 * f) * Kp + e]);
        const Real ny = Real(geo[(G_N0X + 2 * f + 1) * Kp + e]);
        const Real sgn = (isL || boundary) ? Real(1) : Real(-1);
        const Real cnx = sgn * nx, cny = sgn * ny;
        const Real s_own = residual_mode ? Real(len) : Real(len / det);

        // own traces of state and adjoint weight
        Real Uo[NFQ][4], psi[NFQ][4];
#pragma unroll
        for (int q = 0; q < NFQ; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) { Uo[q][v] = ur[v]; psi[q][v] = Real(0); }
#pragma unroll
        for (int q = 0; q < NFQ; ++q)
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                Real ph = tab[O_phiF + (f * NFQ + q) * NP + i];
                Real lf = tab[O_LF + (f * NFQ + q) * NP + i] * (sgn * s_own);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    Uo[q][v] = fma(ph, u[i][v], Uo[q][v]);
                    psi[q][v] = fma(lf, lm[i][v], psi[q][v]);
                }
            }

        Real Un[NFQ][4];
        bool skip = false;
        if (!boundary) {
            const double detn = geo[G_DET * Kp + nb];
            const Real s_nb = -sgn * (residual_mode ? Real(len) : Real(len / detn));
#pragma unroll
            for (int q = 0; q < NFQ; ++q)
#pragma unroll
                for (int v = 0; v < 4; ++v) Un[q][v] = ur[v];
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                Real cu[4], cl[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    cu[v] = uk[sidx(i, v, nb, Kp)];
                    cl[v] = lam[sidx(i, v, nb, Kp)];
                }
                const Real* colp = &s_phiFT[(fn * NP + i) * NFQ];
                const Real* coll = &s_LFT[(fn * NP + i) * NFQ];
#pragma unroll
                for (int q = 0; q < NFQ; ++q) {
                    const int qq = flip ? (NFQ - 1 - q) : q;
                    Real ph = colp[qq], lf = coll[qq] * s_nb;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        Un[q][v] = fma(ph, cu[v], Un[q][v]);
                        psi[q][v] = fma(lf, cl[v], psi[q][v]);
                    }
                }
            }
        } else {
            // ghost states independent of the state: FAR_FIELD conserved, NO_SLIP_WALL from
            // primitives, per face point (the host rejects slip walls, whose ghost depends on
            // the interior state)
            const int type = int(bc_row(sp.bc, nb, NFQ, 0)[0]);
            skip = type < 0;
            const Real gm1 = g - Real(1);
#pragma unroll
            for (int q = 0; q < NFQ; ++q) {
                const double* bcr = bc_row(sp.bc, nb, NFQ, q);
                Real Ug[4] = {Real(bcr[1]), Real(bcr[2]), Real(bcr[3]), Real(bcr[4])};
                if (type == DGX_BC_NO_SLIP_WALL) {
                    const Real r = Ug[0], uu = Ug[1], vv = Ug[2], pp = Ug[3];
                    Ug[1] = r * uu;
                    Ug[2] = r * vv;
                    Ug[3] = pp / gm1 + Real(0.5) * r * (uu * uu + vv * vv);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) Un[q][v] = skip ? Uo[q][v] : Ug[v];
            }
        }
        if (skip) continue;

#pragma unroll
        for (int q = 0; q < NFQ; ++q) {
            Real Aq[4], Bq[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                Aq[v] = (isL || boundary) ? Uo[q][v] : Un[q][v];
                Bq[v] = (isL || boundary) ? Un[q][v] : Uo[q][v];
            }
            Real Ub[4] = {0, 0, 0, 0};
            rusanov_adjoint(Aq, Bq, cnx, cny, psi[q], isL || boundary, g, Ub);
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                Real ph = tab[O_phiF + (f * NFQ + q) * NP + i];
#pragma unroll
                for (int v = 0; v < 4; ++v) ub[i][v] = fma(ph, Ub[v], ub[i][v]);
            }
        }
    }

    // ---------------- output ----------------
    Real* __restrict__ out = static_cast<Real*>(sp.out);
    const Real* __restrict__ add = static_cast<const Real*>(sp.add);
    const Real* __restrict__ acc_in = static_cast<const Real*>(sp.acc_in);
    Real* __restrict__ acc_out = static_cast<Real*>(sp.acc_out);
    if (sp.mode == MODE_RK) {
        const Real dt = Real(sp.dt), cb = Real(sp.cb), ca = Real(sp.ca);
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int k = sidx(i, v, e, Kp);
                Real o = cb * (lm[i][v] + dt * ub[i][v]);
                if (add) o += add[k];
                out[k] = o;
                if (acc_out) acc_out[k] = (acc_in ? acc_in[k] : Real(0)) + ca * lm[i][v];
            }
    } else {
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) out[sidx(i, v, e, Kp)] = ub[i][v];
    }
}

// dJ/du of J = sum_e sum_q w_q |J_e| (m_x^2 + m_y^2) / (2 rho):
// g_i = |J| sum_q w_q phi_q,i (-|m|^2/(2 rho^2), m_x/rho, m_y/rho, 0)
template <typename Real>
__global__ void ke_grad_kernel(const Real* __restrict__ u, const double* __restrict__ tab64,
                               const double* __restrict__ qw, const double* __restrict__ geo,
                               Real* __restrict__ grad, int np, int nq, int phiV_off, int K, int Kp,
                               Off4 uref) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K) return;
    double det = geo[G_DET * Kp + e];
    double gacc[15 * 4];
    for (int k = 0; k < np * 4; ++k) gacc[k] = 0.0;
    for (int q = 0; q < nq; ++q) {
        double U[4] = {uref.v[0], uref.v[1], uref.v[2], uref.v[3]};
        for (int i = 0; i < np; ++i)
            for (int v = 0; v < 4; ++v) U[v] += tab64[phiV_off + q * np + i] * double(u[sidx(i, v, e, Kp)]);
        double ri = 1.0 / U[0];
        double dk[4] = {-0.5 * (U[1] * U[1] + U[2] * U[2]) * ri * ri, U[1] * ri, U[2] * ri, 0.0};
        double w = qw[q] * det;
        for (int i = 0; i < np; ++i) {
            double ph = tab64[phiV_off + q * np + i] * w;
            for (int v = 0; v < 4; ++v) gacc[i * 4 + v] += ph * dk[v];
        }
    }
    for (int k = 0; k < np * 4; ++k) grad[sidx(k >> 2, k & 3, e, Kp)] = Real(gacc[k]);
}

template <typename Real>
hipError_t launch_vjp_real(int p, int rule, const VjpParams& vp, hipStream_t s) {
    dim3 grid((vp.K + kBlock - 1) / kBlock), block(kBlock);
    switch (rule * 8 + p) {
    case 1: hipLaunchKernelGGL((vjp_kernel<1, 0, Real>), grid, block, 0, s, vp); break;
    case 2: hipLaunchKernelGGL((vjp_kernel<2, 0, Real>), grid, block, 0, s, vp); break;
    case 3: hipLaunchKernelGGL((vjp_kernel<3, 0, Real>), grid, block, 0, s, vp); break;
    case 4: hipLaunchKernelGGL((vjp_kernel<4, 0, Real>), grid, block, 0, s, vp); break;
    case 9: hipLaunchKernelGGL((vjp_kernel<1, 1, Real>), grid, block, 0, s, vp); break;
    case 10: hipLaunchKernelGGL((vjp_kernel<2, 1, Real>), grid, block, 0, s, vp); break;
    case 11: hipLaunchKernelGGL((vjp_kernel<3, 1, Real>), grid, block, 0, s, vp); break;
    case 12: hipLaunchKernelGGL((vjp_kernel<4, 1, Real>), grid, block, 0, s, vp); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_vjp(int precision, int p, int rule, const VjpParams& vp, hipStream_t s) {
    if (vp.K <= 0) return hipSuccess;
    if (vp.tabm && vp.conn && precision == DGX_FP32 && rule == 0 && mfma_order(p)) return launch_vjp_mfma(p, vp, s);
    return precision == DGX_FP64 ? launch_vjp_real<double>(p, rule, vp, s) : launch_vjp_real<float>(p, rule, vp, s);
}

hipError_t launch_ke_grad(int precision, int p, int rule, const void* u, const double* tab64, const double* qw,
                          const double* geo, void* grad, int K, int Kp, Off4 uref, hipStream_t s) {
    TabLayout L(p, rule);
    dim3 grid((K + 255) / 256), block(256);
    if (precision == DGX_FP64)
        hipLaunchKernelGGL(ke_grad_kernel<double>, grid, block, 0, s, (const double*)u, tab64, qw, geo, (double*)grad,
                           L.np, L.nq, L.phiV, K, Kp, uref);
    else
        hipLaunchKernelGGL(ke_grad_kernel<float>, grid, block, 0, s, (const float*)u, tab64, qw, geo, (float*)grad,
                           L.np, L.nq, L.phiV, K, Kp, uref);
    return hipGetLastError();
}

}  // namespace dgx
