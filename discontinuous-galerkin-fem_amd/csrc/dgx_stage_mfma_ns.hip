// FP32 Navier-Stokes stage on the matrix cores (gfx950 v_mfma_f32_16x16x4f32), p = 3 and 4,
// volume rule exact to degree 2p.  Same operator and output modes as the per-lane stage kernel
// (dgx_stage_impl.hpp with NS): inviscid and viscous volume terms, Rusanov and viscous face
// terms with the reference's face-gradient read (quirk) or the plain one, inverse mass and the
// SSP-RK3 update.
//
// Reference behaviour (paths relative to the reference root):
//   viscous flux     S/weak_forms/navier_stokes_weak_formulation.cpp:36-98
//   viscous volume   :100-140;  interior faces :142-207;  boundary faces :209-273;  penalty :275-278
//   inviscid terms   S/weak_forms/euler_weak_formulation.cpp:72-250
//   assembly         S/solver/assembler.cpp:137-279
//
// Two launches per stage.  The face-trace kernel evaluates, for every element and every point
// of its three faces, the state and the 8-slot face gradient on the matrix cores, forms the
// viscous flux along the element's outward normal and writes (U, Fv.n) to a record buffer
// rec[face point][element].  The stage kernel then needs no neighbour gradients: it reads the
// two records of each face point (its own and the neighbour's, in the neighbour's point order)
// and combines them as the reference does, with the canonical (L-side) orientation so both
// copies of a face flux come from the same arithmetic.  Normals of the two sides are exact
// negatives (geometry stores the L normal on both), so the neighbour's Fv.n_L is the negative of
// its record.  Volume: U, dU/dxi and dU/deta from the coefficients, and the projection of the
// contravariant inviscid-minus-viscous fluxes, on the matrix cores; the pointwise physics on the
// VALU.  Lane layout as the Euler kernel (dgx_mfma.hpp).
#include "dgx_mfma_dev.hpp"

namespace dgx {
namespace {

// record of one face point: U and the viscous normal flux (first component zero)
struct FaceRecord {
    float4 u, v;
};

template <int P, bool QUIRK>
__device__ __forceinline__ void ns_trace_body(const StageParams& sp) {
    using F = MfmaShape<P>;
    using S = MfmaNsShape<P>;
    constexpr int NP = F::NP, NFQ = F::NFQ, R = F::R;
    constexpr int EB = kBlock / 4;
    __shared__ float s_tab[S::TOTAL_T];

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int e = xcd_tile(blockIdx.x, gridDim.x) * EB + 16 * w + j;
    const bool valid = e < sp.K;
    const int ec = valid ? e : sp.K - 1;
    const unsigned uKp = unsigned(sp.Kp);
    const float gam = float(sp.gamma), mu = float(sp.mu);
    const float kappa = float(sp.kappa);
    const float ur[4] = {float(sp.uref[0]), float(sp.uref[1]), float(sp.uref[2]), float(sp.uref[3])};
    const float* __restrict__ tt = static_cast<const float*>(sp.tabt);

    constexpr int kTabPerThread = (S::TOTAL_T + kBlock - 1) / kBlock;
    float tabv[kTabPerThread];
#pragma unroll
    for (int k = 0; k < kTabPerThread; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        tabv[k] = t < S::TOTAL_T ? tt[t] : 0.0f;
    }
    const unsigned state_bytes = unsigned(NP) * 16u * uKp;
    const auto r_uk = make_rsrc(sp.uk, state_bytes);
    const auto r_geo = make_rsrc(sp.geo, unsigned(kGeoRows) * 8u * uKp);
    auto slot_off = [&](int s, unsigned el) -> unsigned {
        const int m = mf_mode(R, NP, g, s);
        return m >= 0 ? (unsigned(m) * uKp + el) * 16u : state_bytes;
    };
    float b[4][R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const v4u x = load_x4(r_uk, slot_off(s, unsigned(ec)));
#pragma unroll
        for (int v = 0; v < 4; ++v) b[v][s] = __uint_as_float(x[v]);
    }
    const unsigned go = unsigned(ec) * 8u;
    double gA[4], gN[6];
#pragma unroll
    for (int k = 0; k < 4; ++k) gA[k] = load_f64(r_geo, go, unsigned(G_A00 + k) * 8u * uKp);
#pragma unroll
    for (int k = 0; k < 6; ++k) gN[k] = load_f64(r_geo, go, unsigned(G_N0X + k) * 8u * uKp);
#pragma unroll
    for (int k = 0; k < kTabPerThread; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        if (t < S::TOTAL_T) s_tab[t] = tabv[k];
    }
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);

    const float* lt = &s_tab[lane];
    // traces of U at the face slots
    v4f Uf[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < R; ++s) c = mfma16x16x4(lt[(S::T_EF + s) * 64], b[v][s], c);
        Uf[v] = c;
    }
    // face gradients: g[v][c] = sum_{d,k} FG[(c*2+d)*2+k] (A[d][k] u)_v -- the element's
    // dxi_dx entry scales the operand (one column per element), so each (variable, component)
    // accumulates in a single product chain; without the reference's read only the d = c slots
    // are populated
    // Under the reference's read the constant mode has a nonzero gradient row, so its operand
    // is the full coefficient: the stored value plus the storage offset.
    float bg[4][R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const bool m0 = QUIRK && mf_mode(R, NP, g, s) == 0;
#pragma unroll
        for (int v = 0; v < 4; ++v) bg[v][s] = m0 ? b[v][s] + float(sp.uoff[v]) : b[v][s];
    }
    v4f gr[4][2];
#pragma unroll
    for (int v = 0; v < 4; ++v) gr[v][0] = gr[v][1] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int dk = 0; dk < 4; ++dk) {
        const int d = dk >> 1, k = dk & 1;
        const float wA = float(gA[dk]);
        float bs[4][R];
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int s = 0; s < R; ++s) bs[v][s] = wA * bg[v][s];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            if (!QUIRK && d != c) continue;
            const int slot = (c * 2 + d) * 2 + k;
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int s = 0; s < R; ++s) gr[v][c] = mfma16x16x4(lt[(S::T_FG + slot * R + s) * 64], bs[v][s], gr[v][c]);
        }
    }
    if (!valid) return;
    const auto r_rec = make_rsrc(sp.rec, unsigned(3 * NFQ) * 32u * uKp);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int fp = mf_fpoint(NFQ, g, r);
        if (fp < 0) continue;
        const int face = fp / NFQ;
        const float U[4] = {Uf[0][r] + ur[0], Uf[1][r] + ur[1], Uf[2][r] + ur[2], Uf[3][r] + ur[3]};
        float gq[8];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            gq[v * 2] = gr[v][0][r];
            gq[v * 2 + 1] = gr[v][1][r];
        }

        float Fv[4], Gv[4];
        viscous_flux(U, gq, gam, mu, kappa, Fv, Gv);
        const float nx = float(face == 0 ? gN[0] : face == 1 ? gN[2] : gN[4]);
        const float ny = float(face == 0 ? gN[1] : face == 1 ? gN[3] : gN[5]);
        const unsigned ro = (unsigned(fp) * uKp + unsigned(e)) * 32u;
        const v4u xu = {__float_as_uint(U[0]), __float_as_uint(U[1]), __float_as_uint(U[2]), __float_as_uint(U[3])};
        const v4u xv = {0u, __float_as_uint(Fv[1] * nx + Gv[1] * ny), __float_as_uint(Fv[2] * nx + Gv[2] * ny),
                        __float_as_uint(Fv[3] * nx + Gv[3] * ny)};
        __builtin_amdgcn_raw_buffer_store_b128(xu, r_rec, ro, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b128(xv, r_rec, ro + 16u, 0, 0);
    }
}

template <int P>
__device__ __forceinline__ void ns_stage_body(const StageParams& sp) {
    using F = MfmaShape<P>;
    using S = MfmaNsShape<P>;
    constexpr int NP = F::NP, NFQ = F::NFQ, R = F::R, RQ = F::RQ, RF = F::RF;
    constexpr int EB = kBlock / 4;
    __shared__ float s_tab[S::TOTAL_M];
    __shared__ double s_geo[kGeoRows][EB];

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int eb0 = xcd_tile(blockIdx.x, gridDim.x) * EB;
    const int e = eb0 + 16 * w + j;
    const bool valid = e < sp.K;
    const int ec = valid ? e : sp.K - 1;
    const unsigned uKp = unsigned(sp.Kp);
    const bool residual_mode = sp.mode == MODE_RESIDUAL;
    const float gam = float(sp.gamma), mu = float(sp.mu);
    const float kappa = float(sp.kappa);
    const float ur[4] = {float(sp.uref[0]), float(sp.uref[1]), float(sp.uref[2]), float(sp.uref[3])};
    const float* __restrict__ tm = static_cast<const float*>(sp.tabm);
    const auto r_rec = make_rsrc(sp.rec, unsigned(3 * NFQ) * 32u * uKp);
    auto load_record = [&](unsigned idx) {
        FaceRecord r;
        const v4u a = load_x4(r_rec, idx * 32u), b = load_x4(r_rec, idx * 32u + 16u);
        r.u = make_float4(__uint_as_float(a[0]), __uint_as_float(a[1]), __uint_as_float(a[2]), __uint_as_float(a[3]));
        r.v = make_float4(__uint_as_float(b[0]), __uint_as_float(b[1]), __uint_as_float(b[2]), __uint_as_float(b[3]));
        return r;
    };

    // ---------------- loads, issued up front ----------------------------------------------
    constexpr int kTabPerThread = (S::TOTAL_M + kBlock - 1) / kBlock;
    float tabv[kTabPerThread];
#pragma unroll
    for (int k = 0; k < kTabPerThread; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        tabv[k] = t < S::TOTAL_M ? tm[t] : 0.0f;
    }
    const unsigned state_bytes = unsigned(NP) * 16u * uKp;
    const auto r_uk = make_rsrc(sp.uk, state_bytes);
    const auto r_geo = make_rsrc(sp.geo, unsigned(kGeoRows) * 8u * uKp);
    const auto r_nbr = make_rsrc(sp.nbr, 12u * uKp);
    const auto r_code = make_rsrc(sp.code, 4u * uKp);
    const uint32_t code = __builtin_amdgcn_raw_buffer_load_b32(r_code, unsigned(ec) * 4u, 0, 0);
    int nbf[3];
#pragma unroll
    for (int f = 0; f < 3; ++f)
        nbf[f] = int(__builtin_amdgcn_raw_buffer_load_b32(r_nbr, unsigned(ec) * 4u, unsigned(f) * 4u * uKp, 0));
    auto slot_off = [&](int s, unsigned el) -> unsigned {
        const int m = mf_mode(R, NP, g, s);
        return m >= 0 ? (unsigned(m) * uKp + el) * 16u : state_bytes;
    };
    float b[4][R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const v4u x = load_x4(r_uk, slot_off(s, unsigned(ec)));
#pragma unroll
        for (int v = 0; v < 4; ++v) b[v][s] = __uint_as_float(x[v]);
    }
    constexpr int kGeoPerThread = (kGeoRows * EB + kBlock - 1) / kBlock;
    double geov[kGeoPerThread];
#pragma unroll
    for (int k = 0; k < kGeoPerThread; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        const int row = t / EB, el = min(eb0 + t % EB, sp.K - 1);
        geov[k] = t < kGeoRows * EB ? load_f64(r_geo, unsigned(el) * 8u, unsigned(row) * 8u * uKp) : 0.0;
    }
#pragma unroll
    for (int k = 0; k < kTabPerThread; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        if (t < S::TOTAL_M) s_tab[t] = tabv[k];
    }
#pragma unroll
    for (int k = 0; k < kGeoPerThread; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        if (t < kGeoRows * EB) s_geo[t / EB][t % EB] = geov[k];
    }
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);

    const int eloc = 16 * w + j;
    double gd[kGeoRows];
#pragma unroll
    for (int k = 0; k < kGeoRows; ++k) gd[k] = s_geo[k][eloc];
    const float* lt = &s_tab[lane];
    const double det = gd[G_DET];
    const double rdet = 1.0 / det;
    // gradients use dxi_dx; the contravariant fluxes carry |J| in residual mode (linear in dxi_dx)
    const float A00 = float(gd[G_A00]), A01 = float(gd[G_A01]), A10 = float(gd[G_A10]), A11 = float(gd[G_A11]);
    const double sv = residual_mode ? det : 1.0;
    const float a00 = float(sv * gd[G_A00]), a01 = float(sv * gd[G_A01]);
    const float a10 = float(sv * gd[G_A10]), a11 = float(sv * gd[G_A11]);

    // per face: outward normal, lift scale, and the penalty sigma0 mu (p+1)^2 / max(h, 1e-12)
    // (navier_stokes_weak_formulation.cpp:275-278), the latter as an FP32 reciprocal
    const float sig0 = float(sp.sigma0 * sp.mu * (P + 1) * (P + 1));
    float fnx[3], fny[3], fsf[3], fsg[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        fnx[f] = float(gd[G_N0X + 2 * f]);
        fny[f] = float(gd[G_N0X + 2 * f + 1]);
        fsf[f] = residual_mode ? float(gd[G_L0 + f]) : float(gd[G_L0 + f] * rdet);
        fsg[f] = sig0 * rcp(fmaxf(float(gd[G_L0 + f]), 1e-12f));
    }

    // ---------------- volume ----------------------------------------------------------------
    // The projection contracts one register slot per product, so slot r's products follow its
    // fluxes directly: one slot of fluxes is live at a time (same accumulation order per
    // variable as slot-by-slot after all fluxes).
    v4f acc[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    {
        v4f U[4], GX[4], GY[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            v4f cu = {0.0f, 0.0f, 0.0f, 0.0f}, cx = cu, cy = cu;
#pragma unroll
            for (int s = 0; s < R; ++s) {
                cu = mfma16x16x4(lt[(S::M_EV + s) * 64], b[v][s], cu);
                cx = mfma16x16x4(lt[(S::M_GX + s) * 64], b[v][s], cx);
                cy = mfma16x16x4(lt[(S::M_GY + s) * 64], b[v][s], cy);
            }
            U[v] = cu;
            GX[v] = cx;
            GY[v] = cy;
        }
#pragma unroll
        for (int r = 0; r < RQ; ++r) {
            const float Uq[4] = {U[0][r] + ur[0], U[1][r] + ur[1], U[2][r] + ur[2], U[3][r] + ur[3]};
            const float ri = rcp(Uq[0]);
            const float uu = Uq[1] * ri, vv = Uq[2] * ri;
            const float p = (gam - 1.0f) * (Uq[3] - 0.5f * Uq[0] * (uu * uu + vv * vv));
            const float c1 = uu * a00 + vv * a10, c2 = uu * a01 + vv * a11;
            float gq[8];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                gq[v * 2] = A00 * GX[v][r] + A01 * GY[v][r];
                gq[v * 2 + 1] = A10 * GX[v][r] + A11 * GY[v][r];
            }
            float Fv[4], Gv[4];
            viscous_flux(Uq, gq, gam, mu, kappa, Fv, Gv);
            float fh[4], gh[4];
            fh[0] = Uq[0] * c1 - (a00 * Fv[0] + a10 * Gv[0]);
            gh[0] = Uq[0] * c2 - (a01 * Fv[0] + a11 * Gv[0]);
            fh[1] = Uq[1] * c1 + p * a00 - (a00 * Fv[1] + a10 * Gv[1]);
            gh[1] = Uq[1] * c2 + p * a01 - (a01 * Fv[1] + a11 * Gv[1]);
            fh[2] = Uq[2] * c1 + p * a10 - (a00 * Fv[2] + a10 * Gv[2]);
            gh[2] = Uq[2] * c2 + p * a11 - (a01 * Fv[2] + a11 * Gv[2]);
            fh[3] = (Uq[3] + p) * c1 - (a00 * Fv[3] + a10 * Gv[3]);
            gh[3] = (Uq[3] + p) * c2 - (a01 * Fv[3] + a11 * Gv[3]);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                acc[v] = mfma16x16x4(lt[(S::M_PX + r) * 64], fh[v], acc[v]);
                acc[v] = mfma16x16x4(lt[(S::M_PY + r) * 64], gh[v], acc[v]);
            }
        }
    }
    // face records of this lane's slots: its own, and the neighbour's at the matching point
    FaceRecord ro[4], rn[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const int fp = mf_fpoint(NFQ, g, f);
        const int fpc = fp < 0 ? 0 : fp;
        const int face = fpc / NFQ, q = fpc % NFQ;
        ro[f] = load_record(unsigned(fpc) * uKp + unsigned(ec));
        const uint32_t c = (code >> (4 * face)) & 15u;
        const int fn = int(c & 3u);
        const int nb = face == 0 ? nbf[0] : face == 1 ? nbf[1] : nbf[2];
        const bool interior = fn != int(kFaceBoundary) && fp >= 0;
        const int qq = ((c >> 2) & 1u) ? NFQ - 1 - q : q;
        rn[f] = load_record(unsigned(interior ? fn * NFQ + qq : 0) * uKp + unsigned(interior ? nb : ec));
    }

    // ---------------- faces --------------------------------------------------------------
    // Interior faces first, for every slot and without branches (a boundary slot computes a
    // finite value from its own record that the pass after it replaces; a padding slot is
    // selected away); then the boundary slots, only in waves that have one.
    v4f Hs[4];
    bool any_boundary = false;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const int fp = mf_fpoint(NFQ, g, f);
        const bool slot_valid = fp >= 0;
        const int face = slot_valid ? fp / NFQ : 0;
        const uint32_t c = (code >> (4 * face)) & 15u;
        const bool isL = (c >> 3) & 1u;
        any_boundary |= slot_valid && (c & 3u) == kFaceBoundary;
        const float nx = face == 0 ? fnx[0] : face == 1 ? fnx[1] : fnx[2];
        const float ny = face == 0 ? fny[0] : face == 1 ? fny[1] : fny[2];
        const float sf = face == 0 ? fsf[0] : face == 1 ? fsf[1] : fsf[2];
        const float sigma = face == 0 ? fsg[0] : face == 1 ? fsg[1] : fsg[2];
        const float Uo[4] = {ro[f].u.x, ro[f].u.y, ro[f].u.z, ro[f].u.w};
        const float vo[4] = {ro[f].v.x, ro[f].v.y, ro[f].v.z, ro[f].v.w};
        const float Un[4] = {rn[f].u.x, rn[f].u.y, rn[f].u.z, rn[f].u.w};
        const float vn[4] = {rn[f].v.x, rn[f].v.y, rn[f].v.z, rn[f].v.w};
        // canonical orientation: A side = the L element, normal = the L-side outward normal;
        // Fn = 1/2 Fv_L.n -+ 1/2 Fv_R.n - sigma (U_R - U_L) (:192-196; the reference's half
        // difference, or the average) with the own Fv.n_L = sgn vo and the neighbour's =
        // -sgn vn, so both sides form identical numbers
        const float sgn = isL ? 1.0f : -1.0f;
        const float wL = sp.ns_avg_ref ? 0.5f : -0.5f, wR = -0.5f;
        float Aq[4], Bq[4], H[4], Fn[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            Aq[v] = isL ? Uo[v] : Un[v];
            Bq[v] = isL ? Un[v] : Uo[v];
            const float fo = sgn * vo[v], fnb = -sgn * vn[v];
            const float fL = isL ? fo : fnb, fR = isL ? fnb : fo;
            Fn[v] = wL * fL + wR * fR - sigma * (Bq[v] - Aq[v]);
        }
        rusanov(Aq, Bq, sgn * nx, sgn * ny, gam, H);
        const float hs = sgn * sf;
#pragma unroll
        for (int v = 0; v < 4; ++v) Hs[v][f] = slot_valid ? hs * (H[v] + Fn[v]) : 0.0f;
    }
    if (__builtin_amdgcn_ballot_w64(any_boundary) != 0) {
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const int fp = mf_fpoint(NFQ, g, f);
            if (fp < 0) continue;
            const int face = fp / NFQ;
            const uint32_t c = (code >> (4 * face)) & 15u;
            if ((c & 3u) != kFaceBoundary) continue;
            const int nb = face == 0 ? nbf[0] : face == 1 ? nbf[1] : nbf[2];
            const double* bcr = bc_row(sp.bc, nb, NFQ, fp % NFQ);
            if (bcr[0] < 0.0) {  // no condition: the face contributes nothing
#pragma unroll
                for (int v = 0; v < 4; ++v) Hs[v][f] = 0.0f;
                continue;
            }
            const float nx = face == 0 ? fnx[0] : face == 1 ? fnx[1] : fnx[2];
            const float ny = face == 0 ? fny[0] : face == 1 ? fny[1] : fny[2];
            const float sf = face == 0 ? fsf[0] : face == 1 ? fsf[1] : fsf[2];
            const float sigma = face == 0 ? fsg[0] : face == 1 ? fsg[1] : fsg[2];
            const float Uo[4] = {ro[f].u.x, ro[f].u.y, ro[f].u.z, ro[f].u.w};
            const float vo[4] = {ro[f].v.x, ro[f].v.y, ro[f].v.z, ro[f].v.w};
            // ghost state and viscous boundary flux Fv.n - sigma (U_bc - U) (:249-264), U_bc the
            // condition's data at the point itself, taken as conserved except for NO_SLIP
            float Un[4], Ub[4], H[4];
            ghost_state(Uo, nx, ny, bcr, gam, Un);
            if (int(bcr[0]) == DGX_BC_NO_SLIP_WALL) {
                const float Wb[4] = {float(bcr[5]), float(bcr[6]), float(bcr[7]), float(bcr[8])};
                prim2cons(Wb, gam, Ub);
            } else {
#pragma unroll
                for (int v = 0; v < 4; ++v) Ub[v] = float(bcr[5 + v]);
            }
            rusanov(Uo, Un, nx, ny, gam, H);
            const float wB = sp.ns_avg_ref ? 1.0f : -1.0f;
#pragma unroll
            for (int v = 0; v < 4; ++v) Hs[v][f] = sf * (H[v] + (wB * vo[v] - sigma * (Ub[v] - Uo[v])));
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int t = 0; t < RF; ++t) acc[v] = mfma16x16x4(lt[(S::M_LF + t) * 64], Hs[v][t], acc[v]);

    // ---------------- output ---------------------------------------------------------------
    if (!valid) return;
    const auto r_out = make_rsrc(sp.out, state_bytes);
    const float dt = float(sp.dt), ca = float(sp.ca), cb = float(sp.cb);
    const bool rk_blend = sp.mode == MODE_RK && sp.ca != 0.0;
    float unv[4][R];
    if (rk_blend) {
        const auto r_un = make_rsrc(sp.un, state_bytes);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const v4u x = load_x4(r_un, slot_off(r, unsigned(e)));
#pragma unroll
            for (int v = 0; v < 4; ++v) unv[v][r] = __uint_as_float(x[v]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        v4u x;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            float o;
            if (sp.mode != MODE_RK) o = acc[v][r];
            else if (!rk_blend) o = b[v][r] + dt * acc[v][r];
            else o = ca * unv[v][r] + cb * (b[v][r] + dt * acc[v][r]);
            x[v] = __float_as_uint(o);
        }
        __builtin_amdgcn_raw_buffer_store_b128(x, r_out, slot_off(r, unsigned(e)), 0, 0);
    }
}

// four waves per SIMD fit the face-trace kernels without spilling
#define DGX_TRACE_KERNEL(NAME, P, QUIRK)                                                                  \
    __global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void NAME(StageParams sp) { \
        ns_trace_body<P, QUIRK>(sp);                                                                        \
    }
DGX_TRACE_KERNEL(ns_trace_kernel_p3, 3, false)
DGX_TRACE_KERNEL(ns_trace_kernel_p4, 4, false)
DGX_TRACE_KERNEL(ns_trace_quirk_kernel_p3, 3, true)
DGX_TRACE_KERNEL(ns_trace_quirk_kernel_p4, 4, true)
#undef DGX_TRACE_KERNEL
// the stage kernels fit three waves per SIMD without spilling when asked to (the default
// schedule keeps more registers and runs two)
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) void ns_stage_mfma_kernel_p3(StageParams sp) {
    ns_stage_body<3>(sp);
}
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) void ns_stage_mfma_kernel_p4(StageParams sp) {
    ns_stage_body<4>(sp);
}

}  // namespace

hipError_t launch_stage_mfma_ns(int p, bool quirk, const StageParams& sp, hipStream_t s) {
    if (sp.K <= 0) return hipSuccess;
    dim3 grid((sp.K + 63) / 64), block(kBlock);
    switch (p * 2 + (quirk ? 1 : 0)) {
    case 6: hipLaunchKernelGGL(ns_trace_kernel_p3, grid, block, 0, s, sp); break;
    case 7: hipLaunchKernelGGL(ns_trace_quirk_kernel_p3, grid, block, 0, s, sp); break;
    case 8: hipLaunchKernelGGL(ns_trace_kernel_p4, grid, block, 0, s, sp); break;
    case 9: hipLaunchKernelGGL(ns_trace_quirk_kernel_p4, grid, block, 0, s, sp); break;
    default: return hipErrorInvalidValue;
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    if (p == 3) hipLaunchKernelGGL(ns_stage_mfma_kernel_p3, grid, block, 0, s, sp);
    else hipLaunchKernelGGL(ns_stage_mfma_kernel_p4, grid, block, 0, s, sp);
    return hipGetLastError();
}

}  // namespace dgx
