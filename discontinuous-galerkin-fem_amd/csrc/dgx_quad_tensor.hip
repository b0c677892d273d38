// Quadrilateral elements, Euler, p <= 3: the fused DG stage with four lanes per element.
//
// The reference's quad space (Legendre tensor basis phi_{b R + a} = P_a(xi) P_b(eta), tensor
// Gauss-Legendre rule with R = p + 1 nodes per direction, bilinear mapping; S/core/space.cpp:
// 47-55, S/basis/legendre.cpp:69-112) makes every contraction of the stage a product of two
// R x R one-dimensional ones.  A quad of lanes (one DPP row of four) holds one element:
//   * lane a owns the modes (a, 0..p) (xi degree a) and the volume points (s_a, s_0..p);
//   * volume: T_a(eta_q) = sum_b P_b(s_q) u_ab in each lane, U(s_a, s_q) = sum_m P_m(s_a) T_m
//     with the other lanes' T read through quad broadcasts; the projection runs the same two
//     steps backwards;
//   * lane f owns face f: its own trace (the other rows by broadcast), the neighbour's trace
//     (the neighbour's modes streamed row by row from memory), the Rusanov flux; the lift
//     sends each face flux to the four rows by broadcast;
//   * geometry is recomputed from the four vertices: the bilinear map's six derivative
//     coefficients in FP64, dxi/dx and w|J| at the points in the state precision, so per
//     element the stage reads 64 B of vertices and 96 B of face normals and lengths instead
//     of the per-point rows;
//   * the inverse mass: on a parallelogram (code bit kQuadAffineBit) M = |J| diag(m_a m_b), so
//     M^-1 is a scaling of each lane's own modes; otherwise the per-element FP64 inverse, with
//     the residual gathered from the quad by broadcast.
// Reference behaviour as the one-element-per-lane quad kernel: volume and face terms
// S/weak_forms/euler_weak_formulation.cpp:72-250, assembly S/solver/assembler.cpp:137-279,
// M^-1 and SSP-RK3 S/solver/compressible_solver_base.cpp:41-112.
#include "dgx_stage_impl.hpp"

namespace dgx {
namespace {

constexpr int kTBlock = 256;  // 64 elements per block

// The value x holds in lane `src` of this lane's quad (DPP quad_perm broadcast; every lane of a
// quad is active wherever it is used, so no old value is needed).  src is a constant after
// unrolling; the switch selects the encoded control.
template <int SRC>
__device__ __forceinline__ int qb_i(int x) {
    return __builtin_amdgcn_mov_dpp(x, SRC * 0x55, 0xF, 0xF, true);
}
__device__ __forceinline__ int qb_int(int x, int src) {
    switch (src) {
    case 0: return qb_i<0>(x);
    case 1: return qb_i<1>(x);
    case 2: return qb_i<2>(x);
    default: return qb_i<3>(x);
    }
}
__device__ __forceinline__ float qb(float x, int src) { return __builtin_bit_cast(float, qb_int(__builtin_bit_cast(int, x), src)); }
__device__ __forceinline__ double qb(double x, int src) {
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = uint32_t(qb_int(int(uint32_t(b)), src));
    const uint32_t hi = uint32_t(qb_int(int(uint32_t(b >> 32)), src));
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

// FP32: four waves per SIMD (128 VGPRs, a few values spill; three waves fit without spills and
// time 3% slower); FP64 at the compiler's choice.  The volume residual waits in LDS (16 KB per
// block in FP32) while the faces run, which keeps the face section's traces, fluxes and table
// factors within the register budget.
template <int P, typename Real>
__global__ __launch_bounds__(kTBlock) __attribute__((amdgpu_waves_per_eu(sizeof(Real) == 4 ? 4 : 1)))
void quad_tensor_stage_kernel(StageParams sp) {
    constexpr QuadDims D(P);
    constexpr int R = P + 1, NP = D.np;
    static_assert(R <= 4, "one mode row per lane");
    // the one-dimensional factors in LDS (tab[i] = table entry T1 + i, t64 the FP64 nodes,
    // weights and inverse 1-D masses)
    constexpr int NT = D.total - D.T1;
    __shared__ Real tab[NT];
    __shared__ double t64[3 * R];
    for (int i = threadIdx.x; i < NT; i += kTBlock) tab[i] = static_cast<const Real*>(sp.tab)[D.T1 + i];
    if (threadIdx.x < 3 * R) t64[threadIdx.x] = sp.tab64[D.X1 + threadIdx.x];
    __syncthreads();
    constexpr int T1 = 0, D1 = D.D1 - D.T1, FX = D.FX - D.T1, FY = D.FY - D.T1, LX = D.LX - D.T1;
    constexpr int X1 = 0, W1 = R, IM1 = 2 * R;
    const int l = threadIdx.x & 3;
    const int e = xcd_tile(blockIdx.x, gridDim.x) * (kTBlock / 4) + (threadIdx.x >> 2);
    if (e >= sp.K) return;  // the four lanes of an element leave together
    const Real* __restrict__ uk = static_cast<const Real*>(sp.uk);
    const Real* __restrict__ un = static_cast<const Real*>(sp.un);
    Real* __restrict__ out = static_cast<Real*>(sp.out);
    const double* __restrict__ geo = sp.geo;
    const int Kp = sp.Kp;
    const bool own = l < R;    // lane holds a mode row and a volume column
    const int a = own ? l : 0;  // its xi degree / xi node
    const Real gam = Real(sp.gamma);
    const Real ur[4] = {Real(sp.uref[0]), Real(sp.uref[1]), Real(sp.uref[2]), Real(sp.uref[3])};

    Real u[R][4];
#pragma unroll
    for (int b = 0; b < R; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const Real x = uk[sidx(b * R + a, v, e, Kp)];
            u[b][v] = own ? x : Real(0);
        }
    double inv_det_c;  // 1 / |det J| at the centre (parallelograms: the constant Jacobian)

    // ---------------- volume ---------------------------------------------------------------
    Real acc[R][4];
    {
        Real T[R][4];  // T[q][v] = sum_b P_b(s_q) u_ab
#pragma unroll
        for (int q = 0; q < R; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                Real s = Real(0);
#pragma unroll
                for (int b = 0; b < R; ++b) s = fma(tab[T1 + b * R + q], u[b][v], s);
                T[q][v] = s;
            }
        Real c1[R];  // P_m(s_a): weight of row m's T at this lane's xi node
#pragma unroll
        for (int m = 0; m < R; ++m) c1[m] = own ? tab[T1 + m * R + a] : Real(0);
        // bilinear map: dx/dxi = ex0 + ex1 eta, dx/deta = fx0 + ex1 xi (y alike), coefficients
        // from the vertices in FP64, the per-point Jacobians in the state precision
        Real ex0, ex1, fx0, ey0, ey1, fy0;
        {
            double V[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) V[k] = geo[(size_t)(D.gX + k) * Kp + e];
            const double dex0 = 0.25 * ((V[2] - V[0]) + (V[4] - V[6])), dex1 = 0.25 * ((V[0] - V[2]) + (V[4] - V[6]));
            const double dfx0 = 0.25 * ((V[6] - V[0]) + (V[4] - V[2]));
            const double dey0 = 0.25 * ((V[3] - V[1]) + (V[5] - V[7])), dey1 = 0.25 * ((V[1] - V[3]) + (V[5] - V[7]));
            const double dfy0 = 0.25 * ((V[7] - V[1]) + (V[5] - V[3]));
            inv_det_c = 1.0 / fabs(dex0 * dfy0 - dey0 * dfx0);
            ex0 = Real(dex0), ex1 = Real(dex1), fx0 = Real(dfx0), ey0 = Real(dey0), ey1 = Real(dey1), fy0 = Real(dfy0);
        }
        const Real xa = Real(t64[X1 + a]);
        const double wa = t64[W1 + a];
        const Real j10 = fx0 + ex1 * xa, j11 = fy0 + ey1 * xa;  // d/deta at this lane's xi node
        Real S[R][4], S2[R][4];  // sum_q P_b(s_q) F^(s_a, s_q), sum_q P_b'(s_q) G^(s_a, s_q)
#pragma unroll
        for (int b = 0; b < R; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) S[b][v] = S2[b][v] = Real(0);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            Real U[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                Real s = ur[v];
#pragma unroll
                for (int m = 0; m < R; ++m) s = fma(c1[m], qb(T[q][v], m), s);
                U[v] = s;
            }
            const Real eta = Real(t64[X1 + q]);
            const Real j00 = ex0 + ex1 * eta, j01 = ey0 + ey1 * eta;
            const Real det = j00 * j11 - j01 * j10;
            const Real id = rcp(det);
            const Real a00 = j11 * id, a01 = -j01 * id, a10 = -j10 * id, a11 = j00 * id;
            const Real wd = Real(wa * t64[W1 + q]) * fabs(det);
            const Real w = own ? wd : Real(0);
            const Real ri = own ? rcp(U[0]) : Real(0);
            const Real uu = U[1] * ri, vv = U[2] * ri;
            const Real p = (gam - Real(1)) * (U[3] - Real(0.5) * U[0] * (uu * uu + vv * vv));
            const Real F[4] = {U[1], U[1] * uu + p, U[2] * uu, (U[3] + p) * uu};
            const Real Gf[4] = {U[2], U[1] * vv, U[2] * vv + p, (U[3] + p) * vv};
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const Real Fh = w * (a00 * F[v] + a10 * Gf[v]);
                const Real Gh = w * (a01 * F[v] + a11 * Gf[v]);
#pragma unroll
                for (int b = 0; b < R; ++b) {
                    S[b][v] = fma(tab[T1 + b * R + q], Fh, S[b][v]);
                    S2[b][v] = fma(tab[D1 + b * R + q], Gh, S2[b][v]);
                }
            }
        }
        Real c2[R], c3[R];  // P_a'(s_m), P_a(s_m)
#pragma unroll
        for (int m = 0; m < R; ++m) {
            c2[m] = own ? tab[D1 + a * R + m] : Real(0);
            c3[m] = own ? tab[T1 + a * R + m] : Real(0);
        }
#pragma unroll
        for (int b = 0; b < R; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                Real s = Real(0);
#pragma unroll
                for (int m = 0; m < R; ++m) s = fma(c2[m], qb(S[b][v], m), fma(c3[m], qb(S2[b][v], m), s));
                acc[b][v] = s;
            }
    }

    // the volume residual waits in LDS while the faces run
    __shared__ Real sacc[R * 4][kTBlock];
#pragma unroll
    for (int b = 0; b < R; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) sacc[b * 4 + v][threadIdx.x] = acc[b][v];
    // ---------------- faces: lane l evaluates face l ---------------------------------------
    const uint32_t code = sp.code[e];
    const uint32_t c = quad_code(code, l);
    const int fn = int(c & 3u);
    const bool flip = (c >> 2) & 1u;
    const bool isL = (c >> 3) & 1u;
    const bool boundary = (c >> 4) & 1u;
    const int nb = sp.nbr[l * Kp + e];
    const double* gf = geo + (size_t)(D.gF + 3 * l) * Kp + e;
    const Real nx = Real(gf[0]), ny = Real(gf[Kp]);
    const double len = gf[2 * Kp];

    Real Uo[R][4];
    {
        Real fx[R][R], fy[R][R];
#pragma unroll
        for (int q = 0; q < R; ++q)
#pragma unroll
            for (int k = 0; k < R; ++k) {
                fx[q][k] = tab[FX + (l * R + q) * R + k];
                fy[q][k] = tab[FY + (l * R + q) * R + k];
            }
#pragma unroll
        for (int q = 0; q < R; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) Uo[q][v] = ur[v];
#pragma unroll
        for (int m = 0; m < R; ++m) {
            Real bu[R][4];
#pragma unroll
            for (int b = 0; b < R; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) bu[b][v] = qb(u[b][v], m);
#pragma unroll
            for (int q = 0; q < R; ++q)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    Real s = Real(0);
#pragma unroll
                    for (int b = 0; b < R; ++b) s = fma(fy[q][b], bu[b][v], s);
                    Uo[q][v] = fma(fx[q][m], s, Uo[q][v]);
                }
        }
    }

    Real Un[R][4];
    if (!boundary) {
        Real fy[R][R];  // the neighbour's face fn at its point order
        int fxq[R];     // LDS index of its xi factors at point q
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int qq = flip ? (R - 1 - q) : q;
            fxq[q] = FX + (fn * R + qq) * R;
#pragma unroll
            for (int k = 0; k < R; ++k) fy[q][k] = tab[FY + (fn * R + qq) * R + k];
        }
#pragma unroll
        for (int q = 0; q < R; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) Un[q][v] = ur[v];
        // the neighbour's mode rows one at a time (a runtime loop: one row's loads in flight)
#pragma unroll 1
        for (int m = 0; m < R; ++m) {
            Real row[R][4];
#pragma unroll
            for (int b = 0; b < R; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) row[b][v] = uk[sidx(b * R + m, v, nb, Kp)];
#pragma unroll
            for (int q = 0; q < R; ++q)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    Real s = Real(0);
#pragma unroll
                    for (int b = 0; b < R; ++b) s = fma(fy[q][b], row[b][v], s);
                    Un[q][v] = fma(tab[fxq[q] + m], s, Un[q][v]);
                }
        }
    } else {
        // ghost state from the boundary condition at each face point
        const int type = int(bc_row(sp.bc, nb, R, 0)[0]);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const double* bcr = bc_row(sp.bc, nb, R, q);
            const Real ri = rcp(Uo[q][0]);
            Real W[4];
            W[0] = Uo[q][0];
            W[1] = Uo[q][1] * ri;
            W[2] = Uo[q][2] * ri;
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
    const bool no_bc = boundary && bc_row(sp.bc, nb, R, 0)[0] < 0.0;
    const Real sfe = no_bc ? Real(0) : Real(len);
    const Real sgn = (isL || boundary) ? Real(1) : Real(-1);
    const Real cnx = sgn * nx, cny = sgn * ny;
    Real Hs[R][4];
#pragma unroll
    for (int q = 0; q < R; ++q) {
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

#pragma unroll
    for (int b = 0; b < R; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[b][v] = sacc[b * 4 + v][threadIdx.x];
    // lift: row a gains sum over faces m and points q of -w_q/2 P_a(xi) P_b(eta) Hs_m(q).
    // Faces 0, 2 (eta = -+1): P_b(eta) is the same at every point, so the points are summed
    // first with this row's factor -w_q/2 P_a(xi_q); faces 1, 3 (xi = +-1): P_a(xi) is this
    // row's constant, applied after the eta contraction with -w_q/2 P_b(eta_q) (P_0 = 1, so
    // -w_q/2 is the table's a = 0 entry).
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        Real h[R][4];
#pragma unroll
        for (int q = 0; q < R; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) h[q][v] = qb(Hs[q][v], m);
        if (m % 2 == 0) {
            Real z[4] = {0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const Real lx = own ? tab[LX + (m * R + q) * R + a] : Real(0);
#pragma unroll
                for (int v = 0; v < 4; ++v) z[v] = fma(lx, h[q][v], z[v]);
            }
#pragma unroll
            for (int b = 0; b < R; ++b) {
                const Real fyb = tab[FY + (m * R) * R + b];
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[b][v] = fma(fyb, z[v], acc[b][v]);
            }
        } else {
            const Real pa = own ? tab[FX + (m * R) * R + a] : Real(0);
#pragma unroll
            for (int b = 0; b < R; ++b) {
                Real y[4] = {0, 0, 0, 0};
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const Real c = tab[LX + (m * R + q) * R] * tab[FY + (m * R + q) * R + b];
#pragma unroll
                    for (int v = 0; v < 4; ++v) y[v] = fma(c, h[q][v], y[v]);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[b][v] = fma(pa, y[v], acc[b][v]);
            }
        }
    }

    // ---------------- output: R, or M_e^-1 R (FP64) and the RK update ---------------------
    if (sp.mode == MODE_RESIDUAL) {
        if (own)
#pragma unroll
            for (int b = 0; b < R; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) out[sidx(b * R + a, v, e, Kp)] = acc[b][v];
        return;
    }
    Real r[R][4];
    if (code & kQuadAffineBit) {
        const double sa = t64[IM1 + a] * inv_det_c;
#pragma unroll
        for (int b = 0; b < R; ++b) {
            const Real s = Real(sa * t64[IM1 + b]);
#pragma unroll
            for (int v = 0; v < 4; ++v) r[b][v] = s * acc[b][v];
        }
    } else {
        const double* __restrict__ mi = sp.minve;
        double rr[R][4];
#pragma unroll
        for (int b = 0; b < R; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) rr[b][v] = 0.0;
#pragma unroll
        for (int m = 0; m < R; ++m)
#pragma unroll
            for (int bp = 0; bp < R; ++bp) {
                double av[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) av[v] = double(qb(acc[bp][v], m));
                const int j = bp * R + m;
#pragma unroll
                for (int b = 0; b < R; ++b) {
                    const double mm = mi[(size_t)((b * R + a) * NP + j) * Kp + e];
#pragma unroll
                    for (int v = 0; v < 4; ++v) rr[b][v] = fma(mm, av[v], rr[b][v]);
                }
            }
#pragma unroll
        for (int b = 0; b < R; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) r[b][v] = Real(rr[b][v]);
    }
    if (!own) return;
    const Real dt = Real(sp.dt), ca = Real(sp.ca), cb = Real(sp.cb);
#pragma unroll
    for (int b = 0; b < R; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int ix = sidx(b * R + a, v, e, Kp);
            Real o;
            // the own coefficients again from memory (the element's lines are in cache): holding
            // them in registers through the faces costs more than the reload
            if (sp.mode == MODE_RHS) o = r[b][v];
            else if (sp.ca == 0.0) o = uk[ix] + dt * r[b][v];
            else o = ca * un[ix] + cb * (uk[ix] + dt * r[b][v]);
            out[ix] = o;
        }
}

template <typename Real>
hipError_t tensor_family(int p, const StageParams& sp, hipStream_t s) {
    const dim3 grid((sp.K + kTBlock / 4 - 1) / (kTBlock / 4)), block(kTBlock);
    switch (p) {
    case 1: hipLaunchKernelGGL((quad_tensor_stage_kernel<1, Real>), grid, block, 0, s, sp); break;
    case 2: hipLaunchKernelGGL((quad_tensor_stage_kernel<2, Real>), grid, block, 0, s, sp); break;
    case 3: hipLaunchKernelGGL((quad_tensor_stage_kernel<3, Real>), grid, block, 0, s, sp); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_stage_quad_tensor(int precision, int p, const StageParams& sp, hipStream_t s) {
    if (sp.K <= 0) return hipSuccess;
    if (!sp.tab64 || !sp.minve) return hipErrorInvalidValue;
    return precision == DGX_FP64 ? tensor_family<double>(p, sp, s) : tensor_family<float>(p, sp, s);
}

}  // namespace dgx
