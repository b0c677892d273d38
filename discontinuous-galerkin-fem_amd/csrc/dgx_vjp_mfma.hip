// Matrix-core stage adjoint for FP32 Euler, p = 3 and 4 (gfx950 v_mfma_f32_16x16x4f32): the
// transpose of the matrix-core stage kernel (dgx_stage_mfma.hip), same results as the per-lane
// VJP kernel (dgx_vjp.hip) up to summation order.
//
// With L_e = sum_q PX_q F^(U_q) + PY_q G^(U_q) + sum_f s_ef (len_f/|J_e|) sum_q LF_fq H_f(q):
//   volume   ubar += Phi_V^T [ (dF^/dU)^T (PX lambda) + (dG^/dU)^T (PY lambda) ]   (per point)
//   faces    ubar += Phi_F^T (dH/dU_own)^T psi,
//            psi(q) = s_e (len/|J_e|) (LF lambda_e)(q) + s_n (len/|J_n|) (LF lambda_n)(q')
// The four contractions with lambda or the adjoint fluxes (PX lambda, PY lambda, Phi_V^T, LF
// lambda, Phi_F^T) and the two state evaluations run as 16x16x4 products over a 16-element
// tile (lane layout in dgx_mfma.hpp); the neighbour's state and lift-weighted adjoint traces
// are evaluated per block with one element per lane and handed over through LDS.
//
// RK stage adjoint (SSP_RK::step_rk3 transposed, I/solver/time_stepping.hpp:39-55):
//   out = cb (lambda + dt ubar) + add,  acc_out = acc_in + ca lambda.
#include "dgx_mfma_dev.hpp"
#include "dgx_vjp_math.hpp"

namespace dgx {
namespace {

template <int P>
__device__ __forceinline__ void vjp_mfma_body(const VjpParams& sp) {
    using F = MfmaShape<P>;
    using S = MfmaVjpShape<P>;
    constexpr int NP = F::NP, NFQ = F::NFQ, R = F::R, RQ = F::RQ, RF = F::RF;
    constexpr int EB = kBlock / 4;
    __shared__ float s_nt[2 * S::NT];        // neighbour Phi_F and LF tables
    __shared__ float4 s_tu[3 * NFQ * EB];    // neighbour state traces [face][point][element]
    __shared__ float4 s_tp[3 * NFQ * EB];    // neighbour adjoint-flux weights

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int eb0 = xcd_tile(blockIdx.x, gridDim.x) * EB;
    const int e = eb0 + 16 * w + j;
    const bool valid = e < sp.K;
    const int ec = valid ? e : sp.K - 1;
    const int et = min(eb0 + lane, sp.K - 1);  // neighbour role: face w of block element `lane`
    const unsigned uKp = unsigned(sp.Kp);
    const bool residual_mode = sp.mode == MODE_RESIDUAL;
    const float gam = float(sp.gamma);
    const float ur[4] = {float(sp.uref[0]), float(sp.uref[1]), float(sp.uref[2]), float(sp.uref[3])};
    const float* __restrict__ tm = static_cast<const float*>(sp.tabm);

    // ---------------- loads, all issued up front -----------------------------------------
    constexpr int kNtPerThread = (2 * S::NT + kBlock - 1) / kBlock;
    float ntv[kNtPerThread];
#pragma unroll
    for (int k = 0; k < kNtPerThread; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        ntv[k] = t < 2 * S::NT ? tm[S::T_N * 64 + t] : 0.0f;
    }
    const unsigned state_bytes = unsigned(NP) * 16u * uKp;
    const auto r_uk = make_rsrc(sp.uk, state_bytes);
    const auto r_lam = make_rsrc(sp.lam, state_bytes);
    const auto r_geo = make_rsrc(sp.geo, unsigned(kGeoRows) * 8u * uKp);
    const auto r_nbr = make_rsrc(sp.nbr, 12u * uKp);
    const auto r_code = make_rsrc(sp.code, 4u * uKp);
    const uint32_t code = __builtin_amdgcn_raw_buffer_load_b32(r_code, unsigned(ec) * 4u, 0, 0);
    int nbf[3];
#pragma unroll
    for (int f = 0; f < 3; ++f)
        nbf[f] = int(__builtin_amdgcn_raw_buffer_load_b32(r_nbr, unsigned(ec) * 4u, unsigned(f) * 4u * uKp, 0));
    const int ft = w < 3 ? w : 0;
    const uint32_t ct = (__builtin_amdgcn_raw_buffer_load_b32(r_code, unsigned(et) * 4u, 0, 0) >> (4 * ft)) & 15u;
    const int nbt = int(__builtin_amdgcn_raw_buffer_load_b32(r_nbr, unsigned(et) * 4u, unsigned(ft) * 4u * uKp, 0));

    auto slot_off = [&](int s, unsigned el) -> unsigned {
        const int m = mf_mode(R, NP, g, s);
        return m >= 0 ? (unsigned(m) * uKp + el) * 16u : state_bytes;
    };
    float b[4][R], lb[4][R];  // state and adjoint coefficients of this lane's mode slots
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const v4u x = load_x4(r_uk, slot_off(s, unsigned(ec)));
        const v4u y = load_x4(r_lam, slot_off(s, unsigned(ec)));
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            b[v][s] = __uint_as_float(x[v]);
            lb[v][s] = __uint_as_float(y[v]);
        }
    }
    float tEV[R], tXV[R], tYV[R], tEF[R], tLE[R], tVT[RQ], tFT[RF];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        tEV[s] = tm[(S::T_EV + s) * 64 + lane];
        tXV[s] = tm[(S::T_XV + s) * 64 + lane];
        tYV[s] = tm[(S::T_YV + s) * 64 + lane];
        tEF[s] = tm[(S::T_EF + s) * 64 + lane];
        tLE[s] = tm[(S::T_LE + s) * 64 + lane];
    }
#pragma unroll
    for (int t = 0; t < RQ; ++t) tVT[t] = tm[(S::T_VT + t) * 64 + lane];
#pragma unroll
    for (int t = 0; t < RF; ++t) tFT[t] = tm[(S::T_FT + t) * 64 + lane];
    const unsigned go = unsigned(ec) * 8u;
    double gd[kGeoRows];
#pragma unroll
    for (int k = 0; k < kGeoRows; ++k) gd[k] = load_f64(r_geo, go, unsigned(k) * 8u * uKp);
    // neighbour role: the neighbour's coefficients and |J|, this element's face length and |J|
    const bool t_interior = w < 3 && (ct & 3u) != kFaceBoundary;
    float cn[NP][4], cl[NP][4];
    double len_t = 0.0, det_nt = 1.0;
    if (t_interior) {
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const v4u x = load_x4(r_uk, (unsigned(i) * uKp + unsigned(nbt)) * 16u);
            const v4u y = load_x4(r_lam, (unsigned(i) * uKp + unsigned(nbt)) * 16u);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                cn[i][v] = __uint_as_float(x[v]);
                cl[i][v] = __uint_as_float(y[v]);
            }
        }
        len_t = load_f64(r_geo, unsigned(et) * 8u, unsigned(G_L0 + ft) * 8u * uKp);
        det_nt = load_f64(r_geo, unsigned(nbt) * 8u, unsigned(G_DET) * 8u * uKp);
    }
#pragma unroll
    for (int k = 0; k < kNtPerThread; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        if (t < 2 * S::NT) s_nt[t] = ntv[k];
    }
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);

    // ---------------- neighbour traces (waves 0-2) ----------------------------------------
    if (t_interior) {
        const int fn = int(ct & 3u), flip = int((ct >> 2) & 1u);
        const bool isL = (ct >> 3) & 1u;
        // adjoint-flux weight of the neighbour's lift: -s (len/|J_n|) with s = +1 on the L side
        const float s_nb = (isL ? -1.0f : 1.0f) * float(residual_mode ? len_t : len_t / det_nt);
        const float* TP = &s_nt[(fn * 2 + flip) * NP * NFQ];
        const float* TL = TP + S::NT;
        float Un[NFQ][4], Pn[NFQ][4];
#pragma unroll
        for (int q = 0; q < NFQ; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) { Un[q][v] = ur[v]; Pn[q][v] = 0.0f; }
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int q = 0; q < NFQ; ++q) {
                const float tp = TP[i * NFQ + q], tl = TL[i * NFQ + q];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    Un[q][v] = fma(tp, cn[i][v], Un[q][v]);
                    Pn[q][v] = fma(tl, cl[i][v], Pn[q][v]);
                }
            }
#pragma unroll
        for (int q = 0; q < NFQ; ++q) {
            s_tu[(w * NFQ + q) * EB + lane] = make_float4(Un[q][0], Un[q][1], Un[q][2], Un[q][3]);
            s_tp[(w * NFQ + q) * EB + lane] =
                make_float4(s_nb * Pn[q][0], s_nb * Pn[q][1], s_nb * Pn[q][2], s_nb * Pn[q][3]);
        }
    }

    const double det = gd[G_DET];
    const double rdet = 1.0 / det;
    // residual mode carries |J| in the volume term: fold it into the (linear) flux directions
    const double sv = residual_mode ? det : 1.0;
    const float a00 = float(sv * gd[G_A00]), a01 = float(sv * gd[G_A01]);
    const float a10 = float(sv * gd[G_A10]), a11 = float(sv * gd[G_A11]);
    float fnx[3], fny[3], fsf[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        fnx[f] = float(gd[G_N0X + 2 * f]);
        fny[f] = float(gd[G_N0X + 2 * f + 1]);
        fsf[f] = residual_mode ? float(gd[G_L0 + f]) : float(gd[G_L0 + f] * rdet);
    }

    // ---------------- volume ------------------------------------------------------------
    v4f ub[4];  // (dL/du)^T lambda in the mode slots
    {
        v4f U[4], A[4], B[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            v4f cu = {0.0f, 0.0f, 0.0f, 0.0f}, ca_ = cu, cb_ = cu;
#pragma unroll
            for (int s = 0; s < R; ++s) {
                cu = mfma16x16x4(tEV[s], b[v][s], cu);
                ca_ = mfma16x16x4(tXV[s], lb[v][s], ca_);
                cb_ = mfma16x16x4(tYV[s], lb[v][s], cb_);
            }
            U[v] = cu;
            A[v] = ca_;
            B[v] = cb_;
        }
        v4f Ubar[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (r >= RQ) {
#pragma unroll
                for (int v = 0; v < 4; ++v) Ubar[v][r] = 0.0f;
                continue;
            }
            const float Uq[4] = {U[0][r] + ur[0], U[1][r] + ur[1], U[2][r] + ur[2], U[3][r] + ur[3]};
            const float pa[4] = {A[0][r], A[1][r], A[2][r], A[3][r]};
            const float pb[4] = {B[0][r], B[1][r], B[2][r], B[3][r]};
            float ubq[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            flux_dir_adjoint(Uq, a00, a10, pa, 0.0f, gam, ubq);
            flux_dir_adjoint(Uq, a01, a11, pb, 0.0f, gam, ubq);
#pragma unroll
            for (int v = 0; v < 4; ++v) Ubar[v][r] = ubq[v];
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int t = 0; t < RQ; ++t) c = mfma16x16x4(tVT[t], Ubar[v][t], c);
            ub[v] = c;
        }
    }

    // ---------------- faces --------------------------------------------------------------
    v4f Uf[4], Lf[4];  // own state traces and own lift weights of lambda at the face slots
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        v4f cu = {0.0f, 0.0f, 0.0f, 0.0f}, cl_ = cu;
#pragma unroll
        for (int s = 0; s < R; ++s) {
            cu = mfma16x16x4(tEF[s], b[v][s], cu);
            cl_ = mfma16x16x4(tLE[s], lb[v][s], cl_);
        }
        Uf[v] = cu;
        Lf[v] = cl_;
    }
    __syncthreads();  // neighbour traces complete
    v4f Hb[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) Hb[v] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    const int eloc = 16 * w + j;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const int face = f < 3 ? f : g;
        const int q = f < 3 ? g : 4;
        if (f == 3 && (NFQ < 5 || g >= 3)) continue;
        if (f < 3 && g >= NFQ) continue;
        const uint32_t c = (code >> (4 * face)) & 15u;
        const bool boundary = (c & 3u) == kFaceBoundary;
        const bool isL = (c >> 3) & 1u;
        const int nb = face == 0 ? nbf[0] : face == 1 ? nbf[1] : nbf[2];
        const float nx = face == 0 ? fnx[0] : face == 1 ? fnx[1] : fnx[2];
        const float ny = face == 0 ? fny[0] : face == 1 ? fny[1] : fny[2];
        const float sf = face == 0 ? fsf[0] : face == 1 ? fsf[1] : fsf[2];
        const bool own_is_a = isL || boundary;
        const float sgn = own_is_a ? 1.0f : -1.0f;
        const float Uo[4] = {Uf[0][f] + ur[0], Uf[1][f] + ur[1], Uf[2][f] + ur[2], Uf[3][f] + ur[3]};
        float psi[4] = {sgn * sf * Lf[0][f], sgn * sf * Lf[1][f], sgn * sf * Lf[2][f], sgn * sf * Lf[3][f]};
        float Un[4];
        if (boundary) {
            // a face without a condition contributes nothing; FAR_FIELD and NO_SLIP_WALL ghost
            // states are constant (slip walls are rejected by the host)
            const double* bcr = bc_row(sp.bc, nb, NFQ, q);
            if (bcr[0] < 0.0) continue;
            ghost_state(Uo, nx, ny, bcr, gam, Un);
        } else {
            const float4 tu = s_tu[(face * NFQ + q) * EB + eloc];
            const float4 tp = s_tp[(face * NFQ + q) * EB + eloc];
            Un[0] = tu.x; Un[1] = tu.y; Un[2] = tu.z; Un[3] = tu.w;
            psi[0] += tp.x; psi[1] += tp.y; psi[2] += tp.z; psi[3] += tp.w;
        }
        float Aq[4], Bq[4], Ub[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            Aq[v] = own_is_a ? Uo[v] : Un[v];
            Bq[v] = own_is_a ? Un[v] : Uo[v];
        }
        rusanov_adjoint(Aq, Bq, sgn * nx, sgn * ny, psi, own_is_a, gam, Ub);
#pragma unroll
        for (int v = 0; v < 4; ++v) Hb[v][f] = Ub[v];
    }
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int t = 0; t < RF; ++t) ub[v] = mfma16x16x4(tFT[t], Hb[v][t], ub[v]);

    // ---------------- output ---------------------------------------------------------------
    if (!valid) return;
    const auto r_out = make_rsrc(sp.out, state_bytes);
    if (sp.mode == MODE_RK) {
        const float dt = float(sp.dt), cb = float(sp.cb), ca = float(sp.ca);
        const auto r_add = make_rsrc(sp.add, sp.add ? state_bytes : 0u);
        const auto r_ain = make_rsrc(sp.acc_in, sp.acc_in ? state_bytes : 0u);
        const auto r_aout = make_rsrc(sp.acc_out, sp.acc_out ? state_bytes : 0u);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const unsigned vo = slot_off(r, unsigned(e));
            v4u ad = {0u, 0u, 0u, 0u}, ai = {0u, 0u, 0u, 0u};  // absent buffers read zero
            if (sp.add) ad = load_x4(r_add, vo);
            if (sp.acc_in) ai = load_x4(r_ain, vo);
            v4u o, a;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                o[v] = __float_as_uint(cb * (lb[v][r] + dt * ub[v][r]) + __uint_as_float(ad[v]));
                a[v] = __float_as_uint(__uint_as_float(ai[v]) + ca * lb[v][r]);
            }
            __builtin_amdgcn_raw_buffer_store_b128(o, r_out, vo, 0, 0);
            if (sp.acc_out) __builtin_amdgcn_raw_buffer_store_b128(a, r_aout, vo, 0, 0);
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            v4u o;
#pragma unroll
            for (int v = 0; v < 4; ++v) o[v] = __float_as_uint(ub[v][r]);
            __builtin_amdgcn_raw_buffer_store_b128(o, r_out, slot_off(r, unsigned(e)), 0, 0);
        }
    }
}

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 8))) void vjp_mfma_kernel_p3(VjpParams sp) { vjp_mfma_body<3>(sp); }
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 8))) void vjp_mfma_kernel_p4(VjpParams sp) { vjp_mfma_body<4>(sp); }

}  // namespace

hipError_t launch_vjp_mfma(int p, const VjpParams& vp, hipStream_t s) {
    if (vp.K <= 0) return hipSuccess;
    dim3 grid((vp.K + 63) / 64), block(kBlock);
    switch (p) {
    case 3: hipLaunchKernelGGL(vjp_mfma_kernel_p3, grid, block, 0, s, vp); break;
    case 4: hipLaunchKernelGGL(vjp_mfma_kernel_p4, grid, block, 0, s, vp); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace dgx
