// Quadrilateral elements: fused DG stage (one element per lane) and the auxiliary kernels
// (projection, inverse mass, reductions) for the reference's quad space -- Legendre tensor
// basis, tensor Gauss-Legendre volume rule, bilinear mapping (S/core/space.cpp:47-55,
// S/basis/legendre.cpp, S/reference/elements.cpp:61-108).
//
// The mapping is bilinear, so the Jacobian varies inside an element.  The stage rebuilds it
// from the element's four vertices: the map's six derivative coefficients in FP64, dxi/dx and
// w_q |J_q| at the volume points and dxi/dx at the face points (the Navier-Stokes face
// gradients, the neighbour's from the neighbour's vertices) in the state precision; the faces'
// normals and lengths are read.  The stage forms the residual R = M dU/dt with unfolded tables
// and applies M_e^-1 in FP64: on a parallelogram M_e = |J| diag(m_a m_b) (the 1-D Legendre
// masses), otherwise the per-element inverse ([Np*Np][Kp], inverted on the host).  The host's
// per-point geometry rows serve the projection and the reductions.
//
// Reference behaviour: volume, face, boundary and viscous terms as for triangles
// (S/weak_forms/euler_weak_formulation.cpp:72-250, navier_stokes_weak_formulation.cpp:100-278),
// assembly S/solver/assembler.cpp:137-279, M^-1 and SSP-RK3
// S/solver/compressible_solver_base.cpp:41-112.  The Navier-Stokes face gradients follow the
// reference's read of dphi_dx_face (quirk) with dxi/dx taken at the point each stored entry
// belongs to.
#include "dgx_stage_impl.hpp"

namespace dgx {
namespace {

// The bilinear map of element e from its vertices (geometry rows gX..gX+7, FP64): dx/dxi =
// ex0 + ex1 eta, dx/deta = fx0 + ex1 xi and y alike -- the host's quad_map regrouped -- with the
// coefficients rounded to the state precision.
template <typename Real>
struct QuadJac {
    Real ex0, ex1, fx0, ey0, ey1, fy0;
    __device__ __forceinline__ void load(const double* __restrict__ geo, int gX, int Kp, int e, double* det_c = nullptr) {
        double V[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) V[k] = geo[(size_t)(gX + k) * Kp + e];
        const double dex0 = 0.25 * ((V[2] - V[0]) + (V[4] - V[6])), dex1 = 0.25 * ((V[0] - V[2]) + (V[4] - V[6]));
        const double dfx0 = 0.25 * ((V[6] - V[0]) + (V[4] - V[2]));
        const double dey0 = 0.25 * ((V[3] - V[1]) + (V[5] - V[7])), dey1 = 0.25 * ((V[1] - V[3]) + (V[5] - V[7]));
        const double dfy0 = 0.25 * ((V[7] - V[1]) + (V[5] - V[3]));
        if (det_c) *det_c = dex0 * dfy0 - dey0 * dfx0;
        ex0 = Real(dex0), ex1 = Real(dex1), fx0 = Real(dfx0), ey0 = Real(dey0), ey1 = Real(dey1), fy0 = Real(dfy0);
    }
    // A = (dxi/dx, deta/dx, dxi/dy, deta/dy) and det J at (xi, eta)
    __device__ __forceinline__ void at(Real xi, Real eta, Real (&A)[4], Real& det) const {
        const Real j00 = ex0 + ex1 * eta, j01 = ey0 + ey1 * eta;  // d(x, y)/dxi
        const Real j10 = fx0 + ex1 * xi, j11 = fy0 + ey1 * xi;    // d(x, y)/deta
        det = j00 * j11 - j01 * j10;
        const Real id = rcp(det);
        A[0] = j11 * id, A[1] = -j01 * id, A[2] = -j10 * id, A[3] = j00 * id;
    }
};

// Reference coordinates of face point s (a 1-D Gauss-Legendre node) of local face f: the edge
// (v_f, v_f+1) traversed counter-clockwise (face_point_quad on the host).
template <typename Real>
__device__ __forceinline__ void quad_face_xy(int f, Real s, Real& xi, Real& eta) {
    xi = f == 0 ? s : f == 1 ? Real(1) : f == 2 ? -s : Real(-1);
    eta = f == 0 ? Real(-1) : f == 1 ? s : f == 2 ? Real(1) : -s;
}

// Physical gradient of U at face point q of local face f of an element with map J:
// g[v*2+c] = sum_i u[i][v] G(q,i,c), G = the reference's read of the stored physical basis
// gradients (entry (q',i',d) of the interleaved storage, dxi/dx at point q').
// The constant mode's entry is nonzero under the reference's read: it takes the full
// coefficient, the stored value plus the storage offset `off`.
template <int P, bool QUIRK, typename Real>
__device__ __forceinline__ void quad_face_grad(const Real (&u)[QuadDims(P).np][4], const Real* __restrict__ tab,
                                               const Real (&x1)[P + 1], const QuadJac<Real>& J, int f, int q,
                                               const Real (&off)[4], Real* g) {
    constexpr QuadDims D(P);
    constexpr int NP = D.np, NFQ = D.nfq;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        Real acc[4] = {0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            int qs = q, is = i, d = c;
            if (QUIRK) {
                const int k = q * NP + i + c * NFQ * NP;
                d = k % 2;
                qs = (k / 2) / NP;
                is = (k / 2) % NP;
            }
            Real xi, eta, A[4], det;
            quad_face_xy(f, x1[qs], xi, eta);
            J.at(xi, eta, A, det);
            const Real a0 = A[d * 2], a1 = A[d * 2 + 1];
            const Real* gr = tab + D.GF + ((f * NFQ + qs) * NP + is) * 2;
            const Real G = a0 * gr[0] + a1 * gr[1];
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[v] = fma(G, (QUIRK && i == 0) ? u[i][v] + off[v] : u[i][v], acc[v]);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) g[v * 2 + c] = acc[v];
    }
}

template <int P, typename Real, bool NS, bool QUIRK>
__global__ __launch_bounds__(kBlock) void quad_stage_kernel(StageParams sp) {
    constexpr QuadDims D(P);
    constexpr int NP = D.np, NQ = D.nq, NFQ = D.nfq;
    const int e = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (e >= sp.K) return;
    const Real* __restrict__ tab = static_cast<const Real*>(sp.tab);
    const Real* __restrict__ uk = static_cast<const Real*>(sp.uk);
    const Real* __restrict__ un = static_cast<const Real*>(sp.un);
    Real* __restrict__ out = static_cast<Real*>(sp.out);
    const double* __restrict__ geo = sp.geo;
    const int Kp = sp.Kp;
    const Real gam = Real(sp.gamma);
    const Real mu = Real(sp.mu);
    const Real kappa = Real(sp.kappa);
    const Real ur[4] = {Real(sp.uref[0]), Real(sp.uref[1]), Real(sp.uref[2]), Real(sp.uref[3])};
    const Real uo[4] = {Real(sp.uoff[0]), Real(sp.uoff[1]), Real(sp.uoff[2]), Real(sp.uoff[3])};
    const double* __restrict__ t64 = sp.tab64;
    constexpr int R = P + 1;
    Real x1[R];  // the 1-D Gauss-Legendre nodes
#pragma unroll
    for (int a = 0; a < R; ++a) x1[a] = Real(t64[D.X1 + a]);
    QuadJac<Real> J;
    double det_c;  // det J at the centre (parallelograms: the constant Jacobian)
    J.load(geo, D.gX, Kp, e, &det_c);

    Real u[NP][4];
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) u[i][v] = uk[sidx(i, v, e, Kp)];
    Real acc[NP][4];
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[i][v] = Real(0);

    // ---------------- volume: R_i += w_q |J_q| (dphi_i/dx F + dphi_i/dy G) ------------------
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        Real U[4] = {ur[0], ur[1], ur[2], ur[3]};
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const Real ph = tab[D.phiV + q * NP + i];
#pragma unroll
            for (int v = 0; v < 4; ++v) U[v] = fma(ph, u[i][v], U[v]);
        }
        // volume point q = (s_(q mod R), s_(q / R)), weight w_(q mod R) w_(q / R)
        Real A[4], det;
        J.at(x1[q % R], x1[q / R], A, det);
        const Real a00 = A[0], a01 = A[1], a10 = A[2], a11 = A[3];
        const Real wd = Real(t64[D.W1 + q % R] * t64[D.W1 + q / R]) * fabs(det);
        const Real ri = rcp(U[0]);
        const Real uu = U[1] * ri, vv = U[2] * ri;
        const Real p = (gam - Real(1)) * (U[3] - Real(0.5) * U[0] * (uu * uu + vv * vv));
        Real F[4] = {U[1], U[1] * uu + p, U[2] * uu, (U[3] + p) * uu};
        Real G[4] = {U[2], U[1] * vv, U[2] * vv + p, (U[3] + p) * vv};
        if constexpr (NS) {
            Real gx[4] = {0, 0, 0, 0}, gy[4] = {0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                const Real dx = tab[D.GXV + q * NP + i], dy = tab[D.GYV + q * NP + i];
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
            for (int v = 0; v < 4; ++v) { F[v] -= Fv[v]; G[v] -= Gv[v]; }
        }
        // contravariant fluxes against the reference gradients
        Real Fh[4], Gh[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            Fh[v] = wd * (a00 * F[v] + a10 * G[v]);
            Gh[v] = wd * (a01 * F[v] + a11 * G[v]);
        }
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const Real gx = tab[D.GXV + q * NP + i], gy = tab[D.GYV + q * NP + i];
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[i][v] = fma(gx, Fh[v], fma(gy, Gh[v], acc[i][v]));
        }
    }

    // ---------------- faces ----------------------------------------------------------------
    const uint32_t code = sp.code[e];
#pragma unroll 1
    for (int f = 0; f < 4; ++f) {
        const uint32_t c = quad_code(code, f);
        const int fn = int(c & 3u);
        const bool flip = (c >> 2) & 1u;
        const bool isL = (c >> 3) & 1u;
        const bool boundary = (c >> 4) & 1u;
        const int nb = sp.nbr[f * Kp + e];
        const double* gf = geo + (size_t)(D.gF + 3 * f) * Kp + e;
        const Real nx = Real(gf[0]), ny = Real(gf[Kp]);
        const double len = gf[2 * Kp];

        Real Uo[NFQ][4], Un[NFQ][4];
#pragma unroll
        for (int q = 0; q < NFQ; ++q) {
#pragma unroll
            for (int v = 0; v < 4; ++v) Uo[q][v] = ur[v];
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                const Real ph = tab[D.phiF + (f * NFQ + q) * NP + i];
#pragma unroll
                for (int v = 0; v < 4; ++v) Uo[q][v] = fma(ph, u[i][v], Uo[q][v]);
            }
        }
        Real un_c[NS ? NP : 1][4];
        if (!boundary) {
#pragma unroll
            for (int q = 0; q < NFQ; ++q)
#pragma unroll
                for (int v = 0; v < 4; ++v) Un[q][v] = ur[v];
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                Real cf[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) cf[v] = uk[sidx(i, v, nb, Kp)];
                if constexpr (NS) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) un_c[i][v] = cf[v];
                }
#pragma unroll
                for (int q = 0; q < NFQ; ++q) {
                    const int qq = flip ? (NFQ - 1 - q) : q;
                    const Real ph = tab[D.phiF + (fn * NFQ + qq) * NP + i];
#pragma unroll
                    for (int v = 0; v < 4; ++v) Un[q][v] = fma(ph, cf[v], Un[q][v]);
                }
            }
        } else {
            // ghost state from the boundary condition at each face point
            const int type = int(bc_row(sp.bc, nb, NFQ, 0)[0]);
#pragma unroll
            for (int q = 0; q < NFQ; ++q) {
                const double* bcr = bc_row(sp.bc, nb, NFQ, q);
                const Real ri = rcp(Uo[q][0]);
                Real W[4];
                W[0] = Uo[q][0]; W[1] = Uo[q][1] * ri; W[2] = Uo[q][2] * ri;
                W[3] = (gam - Real(1)) * (Uo[q][3] - Real(0.5) * W[0] * (W[1] * W[1] + W[2] * W[2]));
                if (type < 0) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) Un[q][v] = Uo[q][v];
                } else if (type == DGX_BC_FAR_FIELD) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) Un[q][v] = Real(bcr[1 + v]);
                } else if (type == DGX_BC_SLIP_WALL) {
                    const Real vn = W[1] * nx + W[2] * ny;
                    const Real Wg[4] = {W[0], W[1] - Real(2) * vn * nx, W[2] - Real(2) * vn * ny, W[3]};
                    prim2cons(Wg, gam, Un[q]);
                } else {
                    const Real Wb[4] = {Real(bcr[1]), Real(bcr[2]), Real(bcr[3]), Real(bcr[4])};
                    prim2cons(Wb, gam, Un[q]);
                }
            }
        }
        const bool no_bc = boundary && bc_row(sp.bc, nb, NFQ, 0)[0] < 0.0;
        const Real sfe = no_bc ? Real(0) : Real(len);
        const Real sgn = (isL || boundary) ? Real(1) : Real(-1);
        const Real cnx = sgn * nx, cny = sgn * ny;
        Real Hs[NFQ][4];
#pragma unroll
        for (int q = 0; q < NFQ; ++q) {
            Real Aq[4], Bq[4], H[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                Aq[v] = (isL || boundary) ? Uo[q][v] : Un[q][v];
                Bq[v] = (isL || boundary) ? Un[q][v] : Uo[q][v];
            }
            rusanov(Aq, Bq, cnx, cny, gam, H);
#pragma unroll
            for (int v = 0; v < 4; ++v) Hs[q][v] = sgn * sfe * H[v];
        }
        if constexpr (NS) {
            QuadJac<Real> Jn;  // the neighbour's map (its vertices are in cache: it was just read)
            if (!boundary) Jn.load(geo, D.gX, Kp, nb);
            const Real sigma = Real(sp.sigma0 * sp.mu * (P + 1) * (P + 1) / fmax(len, 1e-12));
            // (1/2, -1/2): the reference's half difference; (-1/2, -1/2): the consistent -{Fv}.n
            const Real wL = sp.ns_avg_ref ? Real(0.5) : Real(-0.5), wR = Real(-0.5);
            const Real wB = sp.ns_avg_ref ? Real(1) : Real(-1);
#pragma unroll
            for (int q = 0; q < NFQ; ++q) {
                Real go[8], gn[8], Fvo[4], Gvo[4], Fn[4];
                quad_face_grad<P, QUIRK>(u, tab, x1, J, f, q, uo, go);
                viscous_flux(Uo[q], go, gam, mu, kappa, Fvo, Gvo);
                if (!boundary) {
                    const int qq = flip ? (NFQ - 1 - q) : q;
                    Real Fvn[4], Gvn[4];
                    quad_face_grad<P, QUIRK>(un_c, tab, x1, Jn, fn, qq, uo, gn);
                    viscous_flux(Un[q], gn, gam, mu, kappa, Fvn, Gvn);
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const Real fo = Fvo[v] * cnx + Gvo[v] * cny;
                        const Real fnb = Fvn[v] * cnx + Gvn[v] * cny;
                        const Real fL = isL ? fo : fnb, fR = isL ? fnb : fo;
                        const Real UL = isL ? Uo[q][v] : Un[q][v], UR = isL ? Un[q][v] : Uo[q][v];
                        Fn[v] = wL * fL + wR * fR - sigma * (UR - UL);
                    }
                } else {
                    const double* bcr = bc_row(sp.bc, nb, NFQ, q);
                    Real Ub[4];
                    if (int(bcr[0]) == DGX_BC_NO_SLIP_WALL) {
                        const Real Wb[4] = {Real(bcr[5]), Real(bcr[6]), Real(bcr[7]), Real(bcr[8])};
                        prim2cons(Wb, gam, Ub);
                    } else {
#pragma unroll
                        for (int v = 0; v < 4; ++v) Ub[v] = Real(bcr[5 + v]);
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) Fn[v] = wB * (Fvo[v] * nx + Gvo[v] * ny) - sigma * (Ub[v] - Uo[q][v]);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) Hs[q][v] += sgn * sfe * Fn[v];
            }
        }
        // lift: R_i += LF[f][q][i] Hs_q, LF = -w_q/2 phi_i
#pragma unroll
        for (int q = 0; q < NFQ; ++q)
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                const Real l = tab[D.LF + (f * NFQ + q) * NP + i];
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[i][v] = fma(l, Hs[q][v], acc[i][v]);
            }
    }

    // ---------------- output: R, or M_e^-1 R (FP64) and the RK update ---------------------
    if (sp.mode == MODE_RESIDUAL) {
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) out[sidx(i, v, e, Kp)] = acc[i][v];
        return;
    }
    const double* __restrict__ mi = sp.minve;
    const Real dt = Real(sp.dt), ca = Real(sp.ca), cb = Real(sp.cb);
    const bool affine = (code & kQuadAffineBit) != 0;
    const double inv_det_c = 1.0 / fabs(det_c);
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        double r[4] = {0, 0, 0, 0};
        if (affine) {
            // mode i = (i mod R, i / R): M^-1 = diag(1 / (m_a m_b)) / |J|
            const Real s = Real(t64[D.IM1 + i % R] * inv_det_c * t64[D.IM1 + i / R]);
#pragma unroll
            for (int v = 0; v < 4; ++v) r[v] = double(s * acc[i][v]);
        } else {
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                const double m = mi[(size_t)(i * NP + j) * Kp + e];
#pragma unroll
                for (int v = 0; v < 4; ++v) r[v] = fma(m, double(acc[j][v]), r[v]);
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const Real rv = Real(r[v]);
            Real o;
            if (sp.mode == MODE_RHS) o = rv;
            else if (sp.ca == 0.0) o = u[i][v] + dt * rv;
            else o = ca * un[sidx(i, v, e, Kp)] + cb * (u[i][v] + dt * rv);
            out[sidx(i, v, e, Kp)] = o;
        }
    }
}

template <typename Real, bool NS, bool QUIRK>
hipError_t quad_family(int p, const StageParams& sp, hipStream_t s) {
    const dim3 grid((sp.K + kBlock - 1) / kBlock), block(kBlock);
    switch (p) {
    case 1: hipLaunchKernelGGL((quad_stage_kernel<1, Real, NS, QUIRK>), grid, block, 0, s, sp); break;
    case 2: hipLaunchKernelGGL((quad_stage_kernel<2, Real, NS, QUIRK>), grid, block, 0, s, sp); break;
    case 3: hipLaunchKernelGGL((quad_stage_kernel<3, Real, NS, QUIRK>), grid, block, 0, s, sp); break;
    case 4: hipLaunchKernelGGL((quad_stage_kernel<4, Real, NS, QUIRK>), grid, block, 0, s, sp); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------- auxiliary kernels (runtime Np, FP64 arithmetic) ---------------------------
constexpr int kQRed = 256;
constexpr int kQMaxNp = 25;

// u_e = M_e^-1 sum_q w_q |J_q| phi(x_q) U(x_q), less the storage offset on the constant mode
// (project_initial_condition, S/solver/compressible_solver_base.cpp:63-104).
template <typename Real>
__global__ void quad_project_kernel(const double* __restrict__ Uq, Real* __restrict__ u, const double* __restrict__ phiV,
                                    const double* __restrict__ geo, const double* __restrict__ minve, int np, int nq,
                                    int K, int Kp, Off4 off) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K) return;
    double b[kQMaxNp][4];
    for (int i = 0; i < np; ++i)
        for (int v = 0; v < 4; ++v) b[i][v] = 0.0;
    for (int q = 0; q < nq; ++q) {
        const double wd = geo[(size_t)(5 * q + 4) * Kp + e];
        const double* U = Uq + ((size_t)e * nq + q) * 4;
        for (int i = 0; i < np; ++i) {
            const double s = wd * phiV[q * np + i];
            for (int v = 0; v < 4; ++v) b[i][v] += s * U[v];
        }
    }
    for (int i = 0; i < np; ++i)
        for (int v = 0; v < 4; ++v) {
            double s = 0.0;
            for (int j = 0; j < np; ++j) s += minve[(size_t)(i * np + j) * Kp + e] * b[j][v];
            u[sidx(i, v, e, Kp)] = Real(i == 0 ? s - off.v[v] : s);
        }
}

// r <- M_e^-1 r on host-layout data (apply_mass_inv, :54-61)
__global__ void quad_minv_kernel(const double* __restrict__ minve, double* __restrict__ r, int np, int K, int Kp) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K) return;
    double t[kQMaxNp * 4];
    for (int k = 0; k < np * 4; ++k) t[k] = r[(size_t)e * np * 4 + k];
    for (int i = 0; i < np; ++i)
        for (int v = 0; v < 4; ++v) {
            double s = 0.0;
            for (int j = 0; j < np; ++j) s += minve[(size_t)(i * np + j) * Kp + e] * t[j * 4 + v];
            r[(size_t)e * np * 4 + i * 4 + v] = s;
        }
}

// Per-block FP64 reductions with the triangle kernels' output layout:
//   kind 0: sum_q w_q |J_q| U_q (4), kind 1: kinetic energy, kind 2: [max (|u|+c)/h_min,
//   rho min, rho max, #bad points], kind 3: [#NaN state entries, #infinite state entries].
template <typename Real>
__global__ __launch_bounds__(kQRed) void quad_reduce_kernel(int kind, const Real* __restrict__ u,
                                                            const double* __restrict__ phiV,
                                                            const double* __restrict__ geo,
                                                            const double* __restrict__ hmin, int np, int nq, int K,
                                                            int Kp, double gamma, Off4 uref,
                                                            double* __restrict__ partial) {
    __shared__ double sh[4][kQRed];
    const int e = blockIdx.x * kQRed + threadIdx.x;
    double val[4] = {0.0, 0.0, 0.0, 0.0};
    if (kind == 2) { val[1] = INFINITY; val[2] = -INFINITY; }
    if (e < K) {
        if (kind == 3) {
            for (int k = 0; k < np * 4; ++k) {
                const double x = double(u[sidx(k >> 2, k & 3, e, Kp)]);
                if (isnan(x)) val[0] += 1.0;
                else if (isinf(x)) val[1] += 1.0;
            }
        } else {
            double ws = 0.0;
            for (int q = 0; q < nq; ++q) {
                double U[4] = {uref.v[0], uref.v[1], uref.v[2], uref.v[3]};
                for (int i = 0; i < np; ++i) {
                    const double ph = phiV[q * np + i];
                    for (int v = 0; v < 4; ++v) U[v] += ph * double(u[sidx(i, v, e, Kp)]);
                }
                const double wd = geo[(size_t)(5 * q + 4) * Kp + e];
                if (kind == 0) {
                    for (int v = 0; v < 4; ++v) val[v] += wd * U[v];
                } else if (kind == 1) {
                    val[0] += wd * 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0];
                } else {
                    const double rho = U[0];
                    if (isfinite(rho)) { val[1] = fmin(val[1], rho); val[2] = fmax(val[2], rho); }
                    if (!(rho > 0.0) || !isfinite(rho)) { val[3] += 1.0; continue; }
                    const double uu = U[1] / rho, vv = U[2] / rho;
                    const double p = (gamma - 1.0) * (U[3] - rho * 0.5 * (uu * uu + vv * vv));
                    if (!(p > 0.0) || !isfinite(p)) { val[3] += 1.0; continue; }
                    ws = fmax(ws, sqrt(uu * uu + vv * vv) + sqrt(fmax(gamma * p / rho, 0.0)));
                }
            }
            if (kind == 2) val[0] = ws / hmin[e];
        }
    }
    for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = val[k];
    __syncthreads();
    for (int s = kQRed / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 4; ++k) {
                const double a = sh[k][threadIdx.x], b = sh[k][threadIdx.x + s];
                sh[k][threadIdx.x] = (kind == 2 && (k == 0 || k == 2)) ? fmax(a, b) : (kind == 2 && k == 1) ? fmin(a, b) : a + b;
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) partial[blockIdx.x * 4 + k] = sh[k][0];
}

}  // namespace

hipError_t launch_stage_quad(int precision, int physics, int p, bool quirk, const StageParams& sp, hipStream_t s) {
    if (sp.K <= 0) return hipSuccess;
    const bool f64 = precision == DGX_FP64, ns = physics == DGX_NAVIER_STOKES;
    if (f64) {
        if (!ns) return quad_family<double, false, false>(p, sp, s);
        return quirk ? quad_family<double, true, true>(p, sp, s) : quad_family<double, true, false>(p, sp, s);
    }
    if (!ns) return quad_family<float, false, false>(p, sp, s);
    return quirk ? quad_family<float, true, true>(p, sp, s) : quad_family<float, true, false>(p, sp, s);
}

hipError_t launch_project_quad(int precision, int np, int nq, const double* Uq, void* u, const double* phiV,
                               const double* geo, const double* minve, int K, int Kp, Off4 off, hipStream_t s) {
    if (np > kQMaxNp) return hipErrorInvalidValue;
    const dim3 grid((K + 255) / 256), block(256);
    if (precision == DGX_FP64)
        hipLaunchKernelGGL(quad_project_kernel<double>, grid, block, 0, s, Uq, (double*)u, phiV, geo, minve, np, nq, K, Kp, off);
    else
        hipLaunchKernelGGL(quad_project_kernel<float>, grid, block, 0, s, Uq, (float*)u, phiV, geo, minve, np, nq, K, Kp, off);
    return hipGetLastError();
}

hipError_t launch_minv_quad(int np, const double* minve, double* r, int K, int Kp, hipStream_t s) {
    if (np > kQMaxNp) return hipErrorInvalidValue;
    hipLaunchKernelGGL(quad_minv_kernel, dim3((K + 255) / 256), dim3(256), 0, s, minve, r, np, K, Kp);
    return hipGetLastError();
}

hipError_t launch_reduce_quad(int kind, int precision, int np, int nq, const void* u, const double* phiV,
                              const double* geo, const double* hmin, int K, int Kp, double gamma, Off4 uref,
                              double* partial, int nblocks, hipStream_t s) {
    if (precision == DGX_FP64)
        hipLaunchKernelGGL(quad_reduce_kernel<double>, dim3(nblocks), dim3(kQRed), 0, s, kind, (const double*)u, phiV, geo,
                           hmin, np, nq, K, Kp, gamma, uref, partial);
    else
        hipLaunchKernelGGL(quad_reduce_kernel<float>, dim3(nblocks), dim3(kQRed), 0, s, kind, (const float*)u, phiV, geo,
                           hmin, np, nq, K, Kp, gamma, uref, partial);
    return hipGetLastError();
}

}  // namespace dgx
