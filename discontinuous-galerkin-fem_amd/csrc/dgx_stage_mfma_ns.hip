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
// One launch per stage.  A block owns one 64-element tile of the context's element order
// (Morton order of the centroids, so a tile is a compact patch of the mesh), each wave 16 of
// its elements (lane layout in dgx_mfma.hpp), as in the Euler tile kernel (dgx_stage_mfma.hip).
//
// Faces.  The reference's interior face flux in the canonical (L) orientation is
//   Rusanov(U_L, U_R, n_L) + wL Fv_L.n_L + wR Fv_R.n_L - sigma (U_R - U_L)
// with (wL, wR) = (1/2, -1/2) under the reference's viscous average and (-1/2, -1/2) otherwise,
// and on a boundary face Rusanov(U, U_ghost, n) + wB Fv.n - sigma (U_bc - U), wB = 2 wL.  Every
// face point gets one trace record, formed by the element that owns the point and kept in LDS:
// its state U, the normal flux along the owner's outward normal G = F(U).n + c Fv(U, grad U).n
// with the owner's weight c (2 wL on the L side, 2 wR on the R side, 2 wB on a boundary face)
// and the wave speed s = |u.n| + c_sound.  The normals of the two sides are exact negatives, so
// twice the face flux in an element's own orientation is
//   X = (G_own - G_nb) - (max(s_L, s_R) + 2 sigma) (U_nb - U_own),
// the exact negative of the neighbour's X: both copies of an in-tile face flux agree bit for
// bit.  The neighbours' records on the faces that leave the tile are evaluated per lane from the
// neighbour's coefficients, one face point per lane.
//
// A tile runs in four phases; the LDS table region holds the face-trace and external-face
// tables for the first two and the volume tables for the last two:
//   1. records of the external faces' neighbour points
//   2. the tile's own face points: U and the face-gradient sums on the matrix cores, combined
//      with the element's dxi_dx per point (dgx_mfma.hpp), own records
//   3. volume: U, dU/dxi and dU/deta at the volume points and the projection of the
//      contravariant inviscid-minus-viscous fluxes, on the matrix cores
//   4. face fluxes from the records, boundary faces, lift, RK update.
#include "dgx_mfma_dev.hpp"

namespace dgx {
namespace {

constexpr int kTileEl = kBlock / 4;

// Face-point trace record from the state's deviation Ud = U - ur from the point-value storage
// offset ur: the normal flux with the pressure taken relative to pr = (gamma - 1) ur3 and the
// wave speed |u.n| + c from the full pressure.  A constant pressure in the momentum fluxes adds
// p (int grad phi_i - sum_faces int phi_i n) to the residual, zero under the volume and face
// rules used, so the shift changes nothing but the rounding: at low Mach number the pressure
// is many times the dynamic pressure (71 times at M = 0.1), and the momentum fluxes then carry
// the rounding of the variation instead of that of the full pressure.  The energy flux
// (E + p) u.n is (Ud3 + pd + hr) u.n with hr = ur3 + pr = gamma ur3.
__device__ __forceinline__ void trace_record_dev(const float* Ud, const float* ur, float pr, float hr, float nx,
                                                 float ny, float g, float* Fn, float& s) {
    const float r = Ud[0] + ur[0], mx = Ud[1] + ur[1], my = Ud[2] + ur[2];
    const float ri = rcp(r);
    const float u = mx * ri, v = my * ri;
    const float pd = (g - 1.0f) * (Ud[3] - 0.5f * r * (u * u + v * v));
    const float un = u * nx + v * ny;
    s = fabsf(un) + dsqrt(g * (pd + pr) * ri);
    Fn[0] = r * un;
    Fn[1] = mx * un + pd * nx;
    Fn[2] = my * un + pd * ny;
    Fn[3] = (Ud[3] + pd + hr) * un;
}  // elements per block tile: 4 waves x 16

template <int P>
struct NsTile {
    using S = MfmaNsShape<P>;
    static constexpr int NREC = 3 * S::NFQ * kTileEl;  // records of the tile's own face points
    static constexpr int NT1 = S::TOTAL_TH;             // face-trace + external-face tables (floats)
    static constexpr int REGION = NT1 > S::TOTAL_M ? NT1 : S::TOTAL_M;
};

// Trace records of the tile in dynamic LDS: U, G and s of every own face point
// ([face][point][element]), then those of the external faces' neighbour points ([slot][point],
// the reading element's point order): NREC + halo_cap Nfq entries per array.
struct TraceRecs {
    float4* U;
    float4* G;
    float* S;
};

template <int P, bool QUIRK>
__device__ __forceinline__ void ns_tile(const StageParams& sp, int tile, float* __restrict__ tab,
                                        float4* __restrict__ dyn) {
    using F = MfmaShape<P>;
    using S = MfmaNsShape<P>;
    using T = NsTile<P>;
    constexpr int NP = F::NP, NFQ = F::NFQ, R = F::R, RQ = F::RQ, RF = F::RF, HP = S::HP;
    constexpr int EB = kTileEl, NREC = T::NREC;
    const int eb0 = tile * EB;

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int eloc = 16 * w + j;
    const int e = eb0 + eloc;  // element of this lane in the product layout
    const bool valid = e < sp.K;
    const int ec = valid ? e : sp.K - 1;  // lanes past the end compute on a valid element
    const unsigned uKp = unsigned(sp.Kp);
    const bool residual_mode = sp.mode == MODE_RESIDUAL;
    // wave-uniform scalars rounded on the host (lam = viscous_lambda(mu))
    const float gam = sp.f.gamma, mu = sp.f.mu, kappa = sp.f.kappa, lam = sp.f.lam;
    const float* ur = sp.f.uref;
    // the pressure offset of the momentum fluxes and the energy flux's constant (see
    // trace_record_dev)
    const float pr = (gam - 1.0f) * ur[3], hr = ur[3] + pr;
    // storage offset of the constant mode: under the reference's face-gradient read that mode
    // has a populated gradient row, so its operand is the full coefficient
    const float* uo = sp.f.uoff;
    // parity e(c, q) of the populated dxi_dx row of face point q under the reference's read
    auto row_parity = [](int c, int q) { return ((q * NP + c * NFQ * NP) & 1) != 0; };
    const float* __restrict__ tm = static_cast<const float*>(sp.tabm);
    const float* __restrict__ tt = static_cast<const float*>(sp.tabt);
    const int xs = sp.ext_off[tile];     // the tile's external faces
    const int nint = sp.ext_nint[tile];  // ... of which the first nint have a neighbour
    const int nrs = NREC + (sp.halo_cap > 1 ? sp.halo_cap : 1) * NFQ;  // records per array
    const TraceRecs H{dyn, dyn + nrs, reinterpret_cast<float*>(dyn + 2 * nrs)};
    // record weights of the viscous normal flux: L side, R side, boundary face
    const float cL = sp.ns_avg_ref ? 1.0f : -1.0f, cR = -1.0f, cB = 2.0f * cL;

    // ---------------- loads, issued up front ---------------------------------------------------
    // The external-face list entry of this thread's first pair goes first: the neighbour loads
    // depend on it, and the wait for it (the memory counter drains in issue order) then leaves
    // every load issued after it in flight.
    const int npair = nint * NFQ;
    int hp = (int(threadIdx.x) + 64 * (tile & 3)) & (kBlock - 1);
    const int2 hx = reinterpret_cast<const int2*>(sp.halo)[tile * 64 + hp / NFQ];
    __builtin_amdgcn_sched_barrier(0);
    const unsigned state_bytes = unsigned(NP) * 16u * uKp;
    const auto r_uk = make_rsrc(sp.uk, state_bytes);
    const auto r_geo = make_rsrc(sp.geo32, unsigned(kGeo32NsRows) * 4u * uKp);
    const auto r_conn = make_rsrc(sp.conn, 8u * uKp);
    // geometry entry (row, el), the row a wave-uniform offset (geo_at) or a per-lane one
    auto geo_at = [&](int row, unsigned el) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r_geo, el * 4u, unsigned(row) * 4u * uKp, 0));
    };
    auto geo_at_lane = [&](int row, unsigned el) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r_geo, (unsigned(row) * uKp + el) * 4u, 0, 0));
    };
    const v2u cw = __builtin_bit_cast(v2u, __builtin_amdgcn_raw_buffer_load_b64(r_conn, unsigned(ec) * 8u, 0, 0));
    auto slot_off = [&](int s, unsigned el) -> unsigned {
        const int m = mf_mode(R, NP, g, s);
        return m >= 0 ? (unsigned(m) * uKp + el) * 16u : state_bytes;
    };
    float b[4][R];  // coefficients of this lane's mode slots: product operands and RK input
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const v4u x = load_x4(r_uk, slot_off(s, unsigned(ec)));
#pragma unroll
        for (int v = 0; v < 4; ++v) b[v][s] = __uint_as_float(x[v]);
    }
    // face-trace and external-face tables, staged in LDS by cooperative loads (entries past the
    // table read at an offset past the resource: zeros, and no branch before that wait)
    constexpr int kT1 = (T::NT1 + kBlock - 1) / kBlock;
    const auto r_t1 = make_rsrc(tt, unsigned(T::NT1) * 4u);
    float t1v[kT1];
#pragma unroll
    for (int k = 0; k < kT1; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        t1v[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r_t1, unsigned(t) * 4u, 0, 0));
    }
    // External-face points: pair t = (face hk, point hq) of the tile's external faces with a
    // neighbour, one per thread, dealt from wave (tile mod 4) on so that the tiles sharing a CU
    // load its four SIMDs evenly.  The first 256 pairs (256 / Nfq < 64 faces) come from the
    // fixed-stride copy of the tile's first 64 external faces; neighbour coefficients, dxi_dx and
    // the reading element's normal of the thread's first pair are loaded here.
    float cn[NP][4];
    float hA[4], hnx, hny;
    uint32_t hcode;
    auto halo_load = [&](int2 x) {
        hcode = uint32_t(x.y);
        const bool has = x.x >= 0;
        const unsigned en = has ? unsigned(x.x) : 0u;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            // no neighbour: an offset past the resource, a zero read without traffic
            const v4u y = load_x4(r_uk, has ? (unsigned(i) * uKp + en) * 16u : state_bytes);
#pragma unroll
            for (int v = 0; v < 4; ++v) cn[i][v] = __uint_as_float(y[v]);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) hA[k] = geo_at(G32_G00 + k, en);
        // the neighbour's outward normal: the negated normal of the reading element's face
        const int rel = int((hcode >> 5) & 63u), rf = int((hcode >> 3) & 3u);
        hnx = -geo_at_lane(G32_N0X + 2 * rf, unsigned(eb0 + rel));
        hny = -geo_at_lane(G32_N0Y + 2 * rf, unsigned(eb0 + rel));
    };
    __builtin_amdgcn_sched_barrier(0);
    halo_load(hx);
#pragma unroll
    for (int k = 0; k < kT1; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        if (t < T::NT1) tab[t] = t1v[k];
    }
    __syncthreads();
    // keep the scheduler from sinking the loads above into the arithmetic below
    __builtin_amdgcn_sched_barrier(0);

    // ---------------- 1. records of the external faces' neighbour points ---------------------
    for (;;) {
        if (hp < npair) {
            const int hk = hp / NFQ, hq = hp - hk * NFQ;
            const int fn = int(hcode & 3u), flip = int((hcode >> 2) & 1u);
            const bool read_is_l = (hcode >> 11) & 1u;
            const int qq = flip ? NFQ - 1 - hq : hq;  // the point in the neighbour's order
            const float4* ht = reinterpret_cast<const float4*>(&tab[S::TOTAL_T]) + (fn * NFQ + qq) * HP;
            float Un[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // the deviation from the storage offset
#pragma unroll
            for (int i4 = 0; i4 < (NP + 3) / 4; ++i4) {
                const float4 ph = ht[NP + i4];
                const float phv[4] = {ph.x, ph.y, ph.z, ph.w};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if (4 * i4 + c >= NP) continue;
#pragma unroll
                    for (int v = 0; v < 4; ++v) Un[v] = fma(phv[c], cn[4 * i4 + c][v], Un[v]);
                }
            }
            // gradient: component c = sum_i (dxi_dx[d][0] T_ci0 + dxi_dx[d][1] T_ci1) u_i with the
            // populated row d = c, or e(c, qq) xor (i mod 2) under the reference's read (the
            // constant mode then with the storage offset)
            float Ar[2][2][2];  // [c][i mod 2][k]: the dxi_dx row of each component and mode parity
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int par = 0; par < 2; ++par) {
                    const bool d = QUIRK ? (row_parity(c, qq) != (par != 0)) : (c != 0);
                    Ar[c][par][0] = d ? hA[2] : hA[0];
                    Ar[c][par][1] = d ? hA[3] : hA[1];
                }
            float gq[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                const float4 a = ht[i];
                const float w0 = fmaf(Ar[0][i & 1][0], a.x, Ar[0][i & 1][1] * a.y);
                const float w1 = fmaf(Ar[1][i & 1][0], a.z, Ar[1][i & 1][1] * a.w);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const float ui = (QUIRK && i == 0) ? cn[0][v] + uo[v] : cn[i][v];
                    gq[v * 2] = fmaf(w0, ui, gq[v * 2]);
                    gq[v * 2 + 1] = fmaf(w1, ui, gq[v * 2 + 1]);
                }
            }
            float Fv[4], Gv[4], Fn[4], sw;
            const float Uf_[4] = {Un[0] + ur[0], Un[1] + ur[1], Un[2] + ur[2], Un[3] + ur[3]};
            viscous_flux_l(Uf_, gq, gam, mu, lam, kappa, Fv, Gv);
            trace_record_dev(Un, ur, pr, hr, hnx, hny, gam, Fn, sw);
            const float cn_w = read_is_l ? cR : cL;  // the neighbour holds the other side
            const int k = NREC + hk * NFQ + hq;
            H.U[k] = make_float4(Un[0], Un[1], Un[2], Un[3]);
            H.G[k] = make_float4(Fn[0], Fn[1] + cn_w * (Fv[1] * hnx + Gv[1] * hny),
                                 Fn[2] + cn_w * (Fv[2] * hnx + Gv[2] * hny), Fn[3] + cn_w * (Fv[3] * hnx + Gv[3] * hny));
            H.S[k] = sw;
        }
        // more than 256 external-face points (tiles that are not compact patches): the thread's
        // next pair, 256 on
        hp += kBlock;
        if (hp >= npair) break;
        halo_load(reinterpret_cast<const int2*>(sp.ext)[xs + hp / NFQ]);
    }

    // ---------------- 2. own face points ----------------------------------------------------
    // this lane's element geometry as each phase needs it: dxi_dx and the normals here (their
    // latency covered by the trace products below)
    const float A[2][2] = {{geo_at(G32_G00, unsigned(ec)), geo_at(G32_G01, unsigned(ec))},
                           {geo_at(G32_G10, unsigned(ec)), geo_at(G32_G11, unsigned(ec))}};
    float fnx[3], fny[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        fnx[f] = geo_at(G32_N0X + 2 * f, unsigned(ec));
        fny[f] = geo_at(G32_N0Y + 2 * f, unsigned(ec));
    }
    const float* lt = &tab[lane];  // this lane's entries of the region's lane tables
    constexpr int kTM = (S::TOTAL_M + kBlock - 1) / kBlock;
    float tmv[kTM];
    {
        v4f Uf[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < R; ++s) c = mfma16x16x4(lt[(S::T_EF + s) * 64], b[v][s], c);
            Uf[v] = c;
        }
        // face gradients g[v][c] = sum_k sum_i dxi_dx[d][k] T_ck u_i (the populated row d as in
        // dgx_mfma.hpp), the constant mode's operand with the storage offset under the
        // reference's read
        float b0q[4];  // slot 0 (mode 0 in lanes g = 0)
#pragma unroll
        for (int v = 0; v < 4; ++v) b0q[v] = (QUIRK && g == 0) ? b[v][0] + uo[v] : b[v][0];
        auto bq = [&](int v, int s) { return s == 0 ? b0q[v] : b[v][s]; };
        auto tfg = [&](int c, int k, int s) { return lt[(S::T_FG + (c * 2 + k) * R + s) * 64]; };
        float gr[4][2][4];  // [v][c][register]: scalars, so no register tuples need assembling
        if constexpr (!QUIRK) {
            // d = c: both components from the two reference derivatives S_k at the point
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                v4f s0 = {0.0f, 0.0f, 0.0f, 0.0f}, s1 = s0;
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    s0 = mfma16x16x4(tfg(0, 0, s), bq(v, s), s0);
                    s1 = mfma16x16x4(tfg(0, 1, s), bq(v, s), s1);
                }
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int r = 0; r < 4; ++r) gr[v][c][r] = A[c][0] * s0[r] + A[c][1] * s1[r];
            }
        } else if constexpr (R % 2 == 0) {
            // mode R g + s has the parity of s: per (c, k) the even- and odd-mode sums S_e, S_o,
            // then per row (point q) d = e(c, q) for the even modes and 1 - e(c, q) for the odd;
            // registers r < 3 hold point g, register 3 point 4 (a row parity known here)
            float Ae[2][2], Ao[2][2];  // [c][k], rows of point g
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const bool e = row_parity(c, g);
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    Ae[c][k] = e ? A[1][k] : A[0][k];
                    Ao[c][k] = e ? A[0][k] : A[1][k];
                }
            }
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const v4f z = {0.0f, 0.0f, 0.0f, 0.0f};
                    v4f se0 = mfma16x16x4(tfg(c, 0, 0), bq(v, 0), z), se1 = mfma16x16x4(tfg(c, 1, 0), bq(v, 0), z);
                    v4f so0 = mfma16x16x4(tfg(c, 0, 1), bq(v, 1), z), so1 = mfma16x16x4(tfg(c, 1, 1), bq(v, 1), z);
#pragma unroll
                    for (int s = 2; s < R; s += 2) {
                        se0 = mfma16x16x4(tfg(c, 0, s), bq(v, s), se0);
                        se1 = mfma16x16x4(tfg(c, 1, s), bq(v, s), se1);
                        so0 = mfma16x16x4(tfg(c, 0, s + 1), bq(v, s + 1), so0);
                        so1 = mfma16x16x4(tfg(c, 1, s + 1), bq(v, s + 1), so1);
                    }
                    const int e4c = (4 * NP + c * NFQ * NP) & 1;  // row parity of point 4
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float ae0 = r < 3 ? Ae[c][0] : A[e4c][0], ae1 = r < 3 ? Ae[c][1] : A[e4c][1];
                        const float ao0 = r < 3 ? Ao[c][0] : A[1 - e4c][0], ao1 = r < 3 ? Ao[c][1] : A[1 - e4c][1];
                        gr[v][c][r] = ae0 * se0[r] + ae1 * se1[r] + ao0 * so0[r] + ao1 * so1[r];
                        // formed here: left to the compiler, the sums would sink to the records
                        // below and keep all four product results of every (v, c) live
                        asm volatile("" : "+v"(gr[v][c][r]));
                    }
                }
        } else {
            // odd slot count: the row parity is zero throughout (checked on the host), so d is
            // the mode's parity and dxi_dx[d][k] goes into the operand
            v4f gs[4][2];
#pragma unroll
            for (int v = 0; v < 4; ++v) gs[v][0] = gs[v][1] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                float bs[4][R];
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    const int m = mf_mode(R, NP, g, s);
                    const float a = (m & 1) ? A[1][k] : A[0][k];
#pragma unroll
                    for (int v = 0; v < 4; ++v) bs[v][s] = a * bq(v, s);
                }
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
#pragma unroll
                        for (int s = 0; s < R; ++s) gs[v][c] = mfma16x16x4(tfg(c, k, s), bs[v][s], gs[v][c]);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int r = 0; r < 4; ++r) gr[v][c][r] = gs[v][c][r];
        }
        // volume tables, loaded now and staged once every wave is done with the region
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < kTM; ++k) {
            const int t = int(threadIdx.x) + k * kBlock;
            tmv[k] = t < S::TOTAL_M ? tm[t] : 0.0f;
        }
        // records of the own face points: slot (g, f) = face f, point g (f < 3); face g, point 4
        // (f = 3, p = 4)
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const int fp = mf_fpoint(NFQ, g, f);
            if (f == 3 && NFQ < 5) continue;
            const int face = fp < 0 ? 0 : fp / NFQ, q = fp < 0 ? 0 : fp % NFQ;
            const float Ud[4] = {Uf[0][f], Uf[1][f], Uf[2][f], Uf[3][f]};  // deviation
            const float U[4] = {Ud[0] + ur[0], Ud[1] + ur[1], Ud[2] + ur[2], Ud[3] + ur[3]};
            float gq[8];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                gq[v * 2] = gr[v][0][f];
                gq[v * 2 + 1] = gr[v][1][f];
            }
            const uint32_t c = conn_field(cw[0], cw[1], face);
            const float wv = (c & kConnBoundary) ? cB : ((c >> 13) & 1u) ? cL : cR;
            const float nx = face == 0 ? fnx[0] : face == 1 ? fnx[1] : fnx[2];
            const float ny = face == 0 ? fny[0] : face == 1 ? fny[1] : fny[2];
            float Fv[4], Gv[4], Fn[4], sw;
            viscous_flux_l(U, gq, gam, mu, lam, kappa, Fv, Gv);
            trace_record_dev(Ud, ur, pr, hr, nx, ny, gam, Fn, sw);
            if (fp >= 0) {
                const int k = (face * NFQ + q) * EB + eloc;
                H.U[k] = make_float4(Ud[0], Ud[1], Ud[2], Ud[3]);
                H.G[k] = make_float4(Fn[0], Fn[1] + wv * (Fv[1] * nx + Gv[1] * ny), Fn[2] + wv * (Fv[2] * nx + Gv[2] * ny),
                                     Fn[3] + wv * (Fv[3] * nx + Gv[3] * ny));
                H.S[k] = sw;
            }
        }
    }
    // the rest of the geometry, in flight over the barrier
    const float a00 = geo_at(G32_A00, unsigned(ec)), a01 = geo_at(G32_A01, unsigned(ec));
    const float a10 = geo_at(G32_A10, unsigned(ec)), a11 = geo_at(G32_A11, unsigned(ec));
    float fsf[3], flen[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        fsf[f] = geo_at(G32_S0 + f, unsigned(ec));
        flen[f] = geo_at(G32_L0 + f, unsigned(ec));
    }
    __syncthreads();  // every record of the tile complete, the region's tables done with
#pragma unroll
    for (int k = 0; k < kTM; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        if (t < S::TOTAL_M) tab[t] = tmv[k];
    }
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);

    // ---------------- 3. volume ------------------------------------------------------------
    // gradients use dxi_dx; the contravariant fluxes carry |J| in residual mode
    const float A00 = A[0][0], A01 = A[0][1], A10 = A[1][0], A11 = A[1][1];
    // The projection contracts one register slot per product, so slot r's products follow its
    // fluxes directly: one slot of fluxes is live at a time.
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
            // the pressure relative to pr and the total energy flux's (E + p) from the deviation
            // (see trace_record_dev)
            const float p = (gam - 1.0f) * (U[3][r] - 0.5f * Uq[0] * (uu * uu + vv * vv));
            const float ep = U[3][r] + p + hr;
            const float c1 = uu * a00 + vv * a10, c2 = uu * a01 + vv * a11;
            float gq[8];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                gq[v * 2] = A00 * GX[v][r] + A01 * GY[v][r];
                gq[v * 2 + 1] = A10 * GX[v][r] + A11 * GY[v][r];
            }
            float Fv[4], Gv[4];
            viscous_flux_l(Uq, gq, gam, mu, lam, kappa, Fv, Gv);
            float fh[4], gh[4];
            fh[0] = Uq[0] * c1 - (a00 * Fv[0] + a10 * Gv[0]);
            gh[0] = Uq[0] * c2 - (a01 * Fv[0] + a11 * Gv[0]);
            fh[1] = Uq[1] * c1 + p * a00 - (a00 * Fv[1] + a10 * Gv[1]);
            gh[1] = Uq[1] * c2 + p * a01 - (a01 * Fv[1] + a11 * Gv[1]);
            fh[2] = Uq[2] * c1 + p * a10 - (a00 * Fv[2] + a10 * Gv[2]);
            gh[2] = Uq[2] * c2 + p * a11 - (a01 * Fv[2] + a11 * Gv[2]);
            fh[3] = ep * c1 - (a00 * Fv[3] + a10 * Gv[3]);
            gh[3] = ep * c2 - (a01 * Fv[3] + a11 * Gv[3]);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                acc[v] = mfma16x16x4(lt[(S::M_PX + r) * 64], fh[v], acc[v]);
                acc[v] = mfma16x16x4(lt[(S::M_PY + r) * 64], gh[v], acc[v]);
            }
        }
    }
    // ---------------- 4. faces -------------------------------------------------------------
    // per face: lift scale and twice the penalty 2 sigma0 mu (p+1)^2 / max(h, 1e-12)
    // (navier_stokes_weak_formulation.cpp:275-278), the latter from an FP32 reciprocal of the
    // face length (both sides of a face hold the same length)
    const float sig0 = sp.f.sig0;
    float fsg2[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) fsg2[f] = 2.0f * (sig0 * rcp(fmaxf(flen[f], 1e-12f)));
    (void)residual_mode;
    // Interior faces first, for every slot and without branches (a boundary slot computes a
    // value from its own record that the pass after it replaces; a padding slot is selected
    // away); then the boundary slots, only in waves that have one.
    v4f Hs[4];
    bool any_boundary = false;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const bool slot_valid = f < 3 ? g < NFQ : (NFQ == 5 && g < 3);
        if (f == 3 && NFQ < 5) {
#pragma unroll
            for (int v = 0; v < 4; ++v) Hs[v][f] = 0.0f;
            continue;
        }
        const int face = f < 3 ? f : (g < 3 ? g : 0);
        const int q = f < 3 ? (g < NFQ ? g : 0) : 4;
        const uint32_t c = conn_field(cw[0], cw[1], face);
        const bool bnd = (c & kConnBoundary) != 0, isL = (c >> 13) & 1u;
        any_boundary |= slot_valid && bnd;
        const int ko = (face * NFQ + q) * EB + eloc;
        const uint32_t stepc = (c >> 11) & 3u;
        const int step = stepc == 0u ? EB : stepc == 1u ? -EB : 1;
        const int kn = bnd ? ko : int(c & 2047u) + q * step;
        const float4 uo = H.U[ko], go = H.G[ko];
        const float so = H.S[ko];
        const float4 un = H.U[kn], gn = H.G[kn];
        const float sn = H.S[kn];
        const float sw = ref_max(isL ? so : sn, isL ? sn : so) +
                         (face == 0 ? fsg2[0] : face == 1 ? fsg2[1] : fsg2[2]);
        const float hs = 0.5f * (face == 0 ? fsf[0] : face == 1 ? fsf[1] : fsf[2]);
        const float X0 = (go.x - gn.x) - sw * (un.x - uo.x);
        const float X1 = (go.y - gn.y) - sw * (un.y - uo.y);
        const float X2 = (go.z - gn.z) - sw * (un.z - uo.z);
        const float X3 = (go.w - gn.w) - sw * (un.w - uo.w);
        Hs[0][f] = slot_valid ? hs * X0 : 0.0f;
        Hs[1][f] = slot_valid ? hs * X1 : 0.0f;
        Hs[2][f] = slot_valid ? hs * X2 : 0.0f;
        Hs[3][f] = slot_valid ? hs * X3 : 0.0f;
    }
    if (__builtin_amdgcn_ballot_w64(any_boundary) != 0) {
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            if (f == 3 && NFQ < 5) continue;
            const bool slot_valid = f < 3 ? g < NFQ : (NFQ == 5 && g < 3);
            const int face = f < 3 ? f : (g < 3 ? g : 0);
            const uint32_t c = conn_field(cw[0], cw[1], face);
            if (!slot_valid || !(c & kConnBoundary)) continue;
            const int q = f < 3 ? g : 4;
            const int bf = reinterpret_cast<const int2*>(sp.ext)[xs + int(c & 2047u)].x;
            const double* bcr = bc_row(sp.bc, bf, NFQ, q);
            const float nx = face == 0 ? fnx[0] : face == 1 ? fnx[1] : fnx[2];
            const float ny = face == 0 ? fny[0] : face == 1 ? fny[1] : fny[2];
            const float hs = 0.5f * (face == 0 ? fsf[0] : face == 1 ? fsf[1] : fsf[2]);
            if (bcr[0] < 0.0) {
                // no condition: the face contributes nothing, which with the momentum fluxes'
                // pressure taken relative to pr is the flux of -pr alone (twice the face flux
                // of every other face carries -2 pr n)
                Hs[0][f] = 0.0f;
                Hs[1][f] = hs * (-2.0f * pr * nx);
                Hs[2][f] = hs * (-2.0f * pr * ny);
                Hs[3][f] = 0.0f;
                continue;
            }
            const float sg2 = face == 0 ? fsg2[0] : face == 1 ? fsg2[1] : fsg2[2];
            const int ko = (face * NFQ + q) * EB + eloc;
            const float4 uo = H.U[ko], go = H.G[ko];
            const float Uod[4] = {uo.x, uo.y, uo.z, uo.w};  // deviation
            const float Uo[4] = {uo.x + ur[0], uo.y + ur[1], uo.z + ur[2], uo.w + ur[3]};
            const float Go[4] = {go.x, go.y, go.z, go.w};
            // ghost state and the condition's state U_bc at the point (:249-264), taken as
            // conserved except for NO_SLIP
            float Ug[4], Ub[4], Fg[4], sg;
            ghost_state(Uo, nx, ny, bcr, gam, Ug);
            if (int(bcr[0]) == DGX_BC_NO_SLIP_WALL) {
                const float Wb[4] = {float(bcr[5]), float(bcr[6]), float(bcr[7]), float(bcr[8])};
                prim2cons(Wb, gam, Ub);
            } else {
#pragma unroll
                for (int v = 0; v < 4; ++v) Ub[v] = float(bcr[5 + v]);
            }
            float Ugd[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) Ugd[v] = Ug[v] - ur[v];
            trace_record_dev(Ugd, ur, pr, hr, nx, ny, gam, Fg, sg);
            const float s = ref_max(H.S[ko], sg);
#pragma unroll
            for (int v = 0; v < 4; ++v)
                Hs[v][f] = hs * ((Go[v] + Fg[v]) - s * (Ugd[v] - Uod[v]) - sg2 * ((Ub[v] - ur[v]) - Uod[v]));
        }
    }
    // lift: acc += LF^T H
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int t = 0; t < RF; ++t) acc[v] = mfma16x16x4(lt[(S::M_LF + t) * 64], Hs[v][t], acc[v]);

    // ---------------- output ---------------------------------------------------------------
    if (!valid) return;  // (no barrier follows inside the tile)
    const auto r_out = make_rsrc(sp.out, state_bytes);
    const float dt = sp.f.dt, ca = sp.f.ca, cb = sp.f.cb;
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
        // padding slots address past the end and are dropped
        __builtin_amdgcn_raw_buffer_store_b128(x, r_out, slot_off(r, unsigned(e)), 0, 0);
    }
}

// dynamic LDS of the block: the trace records
__device__ __forceinline__ float4* ns_records_lds() {
    extern __shared__ float4 s_ns_rec[];
    return s_ns_rec;
}

template <int P, bool QUIRK>
__device__ __forceinline__ void ns_tile_body(const StageParams& sp) {
    __shared__ float4 s_region[NsTile<P>::REGION / 4];
    ns_tile<P, QUIRK>(sp, xcd_tile(blockIdx.x, gridDim.x), reinterpret_cast<float*>(s_region), ns_records_lds());
}

// three waves per SIMD (the registers of either order; at p = 4 also three blocks per CU by LDS)
#define DGX_NS_TILE_KERNEL(NAME, P, QUIRK, WAVES)                                                             \
    __global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) void NAME(StageParams sp) { \
        ns_tile_body<P, QUIRK>(sp);                                                                          \
    }
DGX_NS_TILE_KERNEL(ns_tile_kernel_p3, 3, false, 3)
DGX_NS_TILE_KERNEL(ns_tile_kernel_p4, 4, false, 3)
DGX_NS_TILE_KERNEL(ns_tile_quirk_kernel_p3, 3, true, 3)
DGX_NS_TILE_KERNEL(ns_tile_quirk_kernel_p4, 4, true, 3)
#undef DGX_NS_TILE_KERNEL

// dynamic LDS: U, G (float4) and s (float) of the tile's own face points and of its external
// faces' neighbour points (at least one face)
size_t ns_records_bytes(int p, int halo_cap) {
    const size_t nfq = size_t(p + 1);
    const size_t n = 3 * nfq * kTileEl + size_t(halo_cap > 1 ? halo_cap : 1) * nfq;
    return n * (2 * sizeof(float4) + sizeof(float));
}

}  // namespace

hipError_t launch_stage_mfma_ns(int p, bool quirk, const StageParams& sp, hipStream_t s) {
    if (sp.K <= 0) return hipSuccess;
    if (!sp.geo32 || !sp.conn || !sp.ext || !sp.ext_off || !sp.ext_nint || !sp.halo) return hipErrorInvalidValue;
    dim3 grid((sp.K + 63) / 64), block(kBlock);
    const size_t lds = ns_records_bytes(p, sp.halo_cap);
    switch (p * 2 + (quirk ? 1 : 0)) {
    case 6: hipLaunchKernelGGL(ns_tile_kernel_p3, grid, block, lds, s, sp); break;
    case 7: hipLaunchKernelGGL(ns_tile_quirk_kernel_p3, grid, block, lds, s, sp); break;
    case 8: hipLaunchKernelGGL(ns_tile_kernel_p4, grid, block, lds, s, sp); break;
    case 9: hipLaunchKernelGGL(ns_tile_quirk_kernel_p4, grid, block, lds, s, sp); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace dgx
