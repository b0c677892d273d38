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
// tile (lane layout in dgx_mfma.hpp).
//
// A block owns one 64-element tile of the context's element order, as the forward tile kernel
// (the same face connections and external-face lists).  Every face point gets one record in
// LDS, formed by the element that owns it: its state trace U and its lift-weighted adjoint
// trace P = (len/|J|) (LF lambda)(q) of that element (and the wave speed |u.n| + c, as the
// forward's records).  With s_e = +1 on the canonical (L) side
// and s_n = -s_e, the face's adjoint flux weight is psi = s_e (P_own - P_nb), the same number
// on both sides.  The neighbours' records on the faces that leave the tile are evaluated per
// lane from the neighbour's coefficients, one face point per thread.
//
// RK stage adjoint (SSP_RK::step_rk3 transposed, I/solver/time_stepping.hpp:39-55):
//   out = cb (lambda + dt ubar) + add,  acc_out = acc_in + ca lambda.
#include "dgx_mfma_dev.hpp"
#include "dgx_vjp_math.hpp"

namespace dgx {
namespace {

constexpr int kTileEl = kBlock / 4;  // elements per block tile: 4 waves x 16

// LAT: the latency-bound form for grids of at most two tiles per CU, where registers are not
// what limits residency: the neighbours' state and adjoint coefficients loaded together up
// front and the external faces' records formed after the volume work, so that no memory round
// trip stands between the tile's start and its arithmetic.
template <int P, bool LAT>
__device__ __forceinline__ void vjp_tile(const VjpParams& sp, int tile, float* __restrict__ tab,
                                         float4* __restrict__ dyn) {
    using F = MfmaShape<P>;
    using S = MfmaVjpShape<P>;
    constexpr int NP = F::NP, NFQ = F::NFQ, R = F::R, RQ = F::RQ, RF = F::RF, HV = S::HV;
    constexpr int EB = kTileEl, NREC = 3 * NFQ * EB;
    const int eb0 = tile * EB;

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int eloc = 16 * w + j;
    const int e = eb0 + eloc;  // element of this lane in the product layout
    const bool valid = e < sp.K;
    const int ec = valid ? e : sp.K - 1;  // lanes past the end compute on a valid element
    const unsigned uKp = unsigned(sp.Kp);
    const float gam = sp.f.gamma;
    const float* ur = sp.f.uref;
    const float* __restrict__ tm = static_cast<const float*>(sp.tabm);
    const int xs = sp.ext_off[tile];     // the tile's external faces
    const int nint = sp.ext_nint[tile];  // ... of which the first nint have a neighbour
    const int nrs = NREC + (sp.halo_cap > 1 ? sp.halo_cap : 1) * NFQ;  // records per array
    float4* __restrict__ RU = dyn;         // state traces
    float4* __restrict__ RP = dyn + nrs;   // lift-weighted adjoint traces
    float* __restrict__ RS = reinterpret_cast<float*>(dyn + 2 * nrs);  // wave speeds |u.n| + c

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
    const auto r_lam = make_rsrc(sp.lam, state_bytes);
    const auto r_geo = make_rsrc(sp.geo32, unsigned(kGeo32NsRows) * 4u * uKp);
    const auto r_conn = make_rsrc(sp.conn, 8u * uKp);
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
    // the tables, staged in LDS by cooperative loads (entries past the table read at an offset
    // past the resource: zeros, and no branch before that wait)
    constexpr int kTabPerThread = (S::TOTAL + kBlock - 1) / kBlock;
    const auto r_tab = make_rsrc(tm, unsigned(S::TOTAL) * 4u);
    float tabv[kTabPerThread];
#pragma unroll
    for (int k = 0; k < kTabPerThread; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        tabv[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r_tab, unsigned(t) * 4u, 0, 0));
    }
    // External-face points: pair t = (face hk, point hq) of the tile's external faces with a
    // neighbour, one per thread, dealt from wave (tile mod 4) on; the first 256 pairs (256 / Nfq
    // < 64 faces) come from the fixed-stride copy of the tile's first 64 external faces.  The
    // neighbour's state coefficients of the thread's first pair are loaded here, its adjoint
    // coefficients once its state trace is formed.
    float cn[NP][4], cl[NP][4];  // the neighbour's state and (LAT) adjoint coefficients
    float hsf;       // the neighbour's lift scale of the shared face
    float hnx, hny;  // the reading element's outward normal of the face
    uint32_t hcode;
    unsigned hen;
    auto halo_load = [&](int2 x) {
        hcode = uint32_t(x.y);
        const bool has = x.x >= 0;
        hen = has ? unsigned(x.x) : 0u;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            // no neighbour: an offset past the resource, a zero read without traffic
            const v4u y = load_x4(r_uk, has ? (unsigned(i) * uKp + hen) * 16u : state_bytes);
#pragma unroll
            for (int v = 0; v < 4; ++v) cn[i][v] = __uint_as_float(y[v]);
            if constexpr (LAT) {
                const v4u z = load_x4(r_lam, has ? (unsigned(i) * uKp + hen) * 16u : state_bytes);
#pragma unroll
                for (int v = 0; v < 4; ++v) cl[i][v] = __uint_as_float(z[v]);
            }
        }
        hsf = geo_at_lane(G32_S0 + int(hcode & 3u), hen);
        const int rel = int((hcode >> 5) & 63u), rf = int((hcode >> 3) & 3u);
        hnx = geo_at_lane(G32_N0X + 2 * rf, unsigned(eb0 + rel));
        hny = geo_at_lane(G32_N0Y + 2 * rf, unsigned(eb0 + rel));
    };
    __builtin_amdgcn_sched_barrier(0);
    halo_load(hx);
    // RK stage adjoint: the accumulator acc_out = acc_in + ca lambda and the output's part that
    // needs no face or volume term, X = cb lambda + add.  At p = 3 the addend and accumulator are
    // loaded here and used at the output, so the tile's first arithmetic does not wait for them
    // (an absent buffer is a resource of size zero over lambda's memory: its loads read zero
    // without traffic).  At p = 4 X and acc_out are formed here, as soon as lambda is in, to
    // keep those registers free.
    const bool rk = sp.mode == MODE_RK;
    const bool has_add = rk && sp.add, has_ain = rk && sp.acc_in;
    const auto r_add = make_rsrc(has_add ? sp.add : sp.lam, has_add ? state_bytes : 0u);
    const auto r_ain = make_rsrc(has_ain ? sp.acc_in : sp.lam, has_ain ? state_bytes : 0u);
    const auto r_aout = make_rsrc(sp.acc_out, sp.acc_out ? state_bytes : 0u);
    // (at p = 4 the deferred form spills registers)
    constexpr bool DEFER = LAT || P == 3;
    float xo[4][R];
    v4u adv[DEFER ? R : 1], aiv[DEFER ? R : 1];
    // X and acc_out of mode slot r from the addend and accumulator words
    auto rk_terms = [&](int r, const v4u& ad, const v4u& ai) {
        const float cb = sp.f.cb, ca = sp.f.ca;
        v4u a;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            xo[v][r] = cb * lb[v][r] + __uint_as_float(ad[v]);
            a[v] = __float_as_uint(__uint_as_float(ai[v]) + ca * lb[v][r]);
        }
        if (sp.acc_out && valid) __builtin_amdgcn_raw_buffer_store_b128(a, r_aout, slot_off(r, unsigned(ec)), 0, 0);
    };
    if constexpr (DEFER) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            adv[r] = load_x4(r_add, slot_off(r, unsigned(ec)));
            aiv[r] = load_x4(r_ain, slot_off(r, unsigned(ec)));
        }
    } else if (rk) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const unsigned vo = slot_off(r, unsigned(ec));
            v4u ad = {0u, 0u, 0u, 0u}, ai = {0u, 0u, 0u, 0u};  // absent buffers read zero
            if (sp.add) ad = load_x4(r_add, vo);
            if (sp.acc_in) ai = load_x4(r_ain, vo);
            rk_terms(r, ad, ai);
        }
    }
#pragma unroll
    for (int k = 0; k < kTabPerThread; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        if (t < S::TOTAL) tab[t] = tabv[k];
    }
    __syncthreads();
    // keep the scheduler from sinking the loads above into the arithmetic below
    __builtin_amdgcn_sched_barrier(0);

    // ---------------- records of the external faces' neighbour points ----------------------
    auto external_records = [&]() {
        for (;;) {
            if (hp < npair) {
                const int hk = hp / NFQ, hq = hp - hk * NFQ;
                const int fn = int(hcode & 3u), flip = int((hcode >> 2) & 1u);
                const int qq = flip ? NFQ - 1 - hq : hq;  // the point in the neighbour's order
                const float4* ht = reinterpret_cast<const float4*>(&tab[S::T_N * 64]) + (fn * NFQ + qq) * HV;
                float Un[4] = {ur[0], ur[1], ur[2], ur[3]};
#pragma unroll
                for (int m = 0; m < HV; ++m) {
                    const float4 t = ht[m];
#pragma unroll
                    for (int v = 0; v < 4; ++v) Un[v] = fma(t.x, cn[2 * m][v], Un[v]);
                    if (2 * m + 1 < NP) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) Un[v] = fma(t.z, cn[2 * m + 1][v], Un[v]);
                    }
                }
                // the neighbour's adjoint coefficients (otherwise loaded here, into the same registers)
                if constexpr (!LAT) {
#pragma unroll
                    for (int i = 0; i < NP; ++i) {
                        const v4u y = load_x4(r_lam, (unsigned(i) * uKp + hen) * 16u);
#pragma unroll
                        for (int v = 0; v < 4; ++v) cn[i][v] = __uint_as_float(y[v]);
                    }
                }
                const auto& ca_n = LAT ? cl : cn;
                float Pn[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int m = 0; m < HV; ++m) {
                    const float4 t = ht[m];
#pragma unroll
                    for (int v = 0; v < 4; ++v) Pn[v] = fma(t.y, ca_n[2 * m][v], Pn[v]);
                    if (2 * m + 1 < NP) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) Pn[v] = fma(t.w, ca_n[2 * m + 1][v], Pn[v]);
                    }
                }
                const int k = NREC + hk * NFQ + hq;
                RU[k] = make_float4(Un[0], Un[1], Un[2], Un[3]);
                RP[k] = make_float4(hsf * Pn[0], hsf * Pn[1], hsf * Pn[2], hsf * Pn[3]);
                RS[k] = wave_speed(Un, hnx, hny, gam);
            }
            // more than 256 external-face points (tiles that are not compact patches): the thread's
            // next pair, 256 on
            hp += kBlock;
            if (hp >= npair) break;
            halo_load(reinterpret_cast<const int2*>(sp.ext)[xs + hp / NFQ]);
        }
    };
    if constexpr (!LAT) external_records();

    // ---------------- own face points -------------------------------------------------------
    // this lane's element geometry (its latency covered by the trace products below)
    float fnx[3], fny[3], fsf[3];  // outward normal and lift scale of each face
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        fsf[f] = geo_at(G32_S0 + f, unsigned(ec));
        fnx[f] = geo_at(G32_N0X + 2 * f, unsigned(ec));
        fny[f] = geo_at(G32_N0Y + 2 * f, unsigned(ec));
    }
    const float a00 = geo_at(G32_A00, unsigned(ec)), a01 = geo_at(G32_A01, unsigned(ec));
    const float a10 = geo_at(G32_A10, unsigned(ec)), a11 = geo_at(G32_A11, unsigned(ec));
    const float* lt = &tab[lane];  // this lane's entries of the lane tables
    {
        v4f Uf[4], Lf[4];  // state traces and lift weights of lambda at the face slots
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            v4f cu = {0.0f, 0.0f, 0.0f, 0.0f}, cl = cu;
#pragma unroll
            for (int s = 0; s < R; ++s) {
                cu = mfma16x16x4(lt[(S::T_EF + s) * 64], b[v][s], cu);
                cl = mfma16x16x4(lt[(S::T_LE + s) * 64], lb[v][s], cl);
            }
            Uf[v] = cu;
            Lf[v] = cl;
        }
        // slot (g, f) = face f, point g (f < 3); face g, point 4 (f = 3, p = 4)
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const int fp = mf_fpoint(NFQ, g, f);
            if (fp < 0) continue;
            const int face = fp / NFQ;
            const float sf = face == 0 ? fsf[0] : face == 1 ? fsf[1] : fsf[2];
            const float nx = face == 0 ? fnx[0] : face == 1 ? fnx[1] : fnx[2];
            const float ny = face == 0 ? fny[0] : face == 1 ? fny[1] : fny[2];
            const float Uo[4] = {Uf[0][f] + ur[0], Uf[1][f] + ur[1], Uf[2][f] + ur[2], Uf[3][f] + ur[3]};
            RU[fp * EB + eloc] = make_float4(Uo[0], Uo[1], Uo[2], Uo[3]);
            RP[fp * EB + eloc] = make_float4(sf * Lf[0][f], sf * Lf[1][f], sf * Lf[2][f], sf * Lf[3][f]);
            RS[fp * EB + eloc] = wave_speed(Uo, nx, ny, gam);
        }
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
                cu = mfma16x16x4(lt[(S::T_EV + s) * 64], b[v][s], cu);
                ca_ = mfma16x16x4(lt[(S::T_XV + s) * 64], lb[v][s], ca_);
                cb_ = mfma16x16x4(lt[(S::T_YV + s) * 64], lb[v][s], cb_);
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
            for (int t = 0; t < RQ; ++t) c = mfma16x16x4(lt[(S::T_VT + t) * 64], Ubar[v][t], c);
            ub[v] = c;
        }
    }
    if constexpr (LAT) external_records();
    __syncthreads();  // every record of the tile complete

    // ---------------- faces --------------------------------------------------------------
    // Interior faces first, for every slot and without branches (a boundary slot reads its own
    // record as the neighbour's, a finite value the pass after it replaces; a padding slot is
    // selected away); then the boundary slots, only in waves that have one.
    v4f Hb[4];
    bool any_boundary = false;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const bool slot_valid = f < 3 ? g < NFQ : (NFQ == 5 && g < 3);
        if (f == 3 && NFQ < 5) {
#pragma unroll
            for (int v = 0; v < 4; ++v) Hb[v][f] = 0.0f;
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
        const float4 uo = RU[ko], po = RP[ko], un = RU[kn], pn = RP[kn];
        const float so = RS[ko], sn = RS[kn];
        const float sgn = isL ? 1.0f : -1.0f;
        const float nx = face == 0 ? fnx[0] : face == 1 ? fnx[1] : fnx[2];
        const float ny = face == 0 ? fny[0] : face == 1 ? fny[1] : fny[2];
        const float Uo[4] = {uo.x, uo.y, uo.z, uo.w}, Un[4] = {un.x, un.y, un.z, un.w};
        const float psi[4] = {sgn * (po.x - pn.x), sgn * (po.y - pn.y), sgn * (po.z - pn.z), sgn * (po.w - pn.w)};
        float Aq[4], Bq[4], Ub[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            Aq[v] = isL ? Uo[v] : Un[v];
            Bq[v] = isL ? Un[v] : Uo[v];
        }
        rusanov_adjoint_s(Aq, Bq, sgn * nx, sgn * ny, psi, isL, isL ? so : sn, isL ? sn : so, gam, Ub);
#pragma unroll
        for (int v = 0; v < 4; ++v) Hb[v][f] = slot_valid ? Ub[v] : 0.0f;
        // one slot's records in registers at a time
        __builtin_amdgcn_sched_barrier(0);
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
            // a face without a condition contributes nothing; FAR_FIELD and NO_SLIP_WALL ghost
            // states are constant (slip walls are rejected by the host)
            if (bcr[0] < 0.0) {
#pragma unroll
                for (int v = 0; v < 4; ++v) Hb[v][f] = 0.0f;
                continue;
            }
            const float nx = face == 0 ? fnx[0] : face == 1 ? fnx[1] : fnx[2];
            const float ny = face == 0 ? fny[0] : face == 1 ? fny[1] : fny[2];
            const int ko = (face * NFQ + q) * EB + eloc;
            const float4 uo = RU[ko], po = RP[ko];
            const float Uo[4] = {uo.x, uo.y, uo.z, uo.w};
            const float psi[4] = {po.x, po.y, po.z, po.w};
            float Un[4], Ub[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            ghost_state(Uo, nx, ny, bcr, gam, Un);
            rusanov_adjoint_s(Uo, Un, nx, ny, psi, true, RS[ko], wave_speed(Un, nx, ny, gam), gam, Ub);
#pragma unroll
            for (int v = 0; v < 4; ++v) Hb[v][f] = Ub[v];
            __builtin_amdgcn_sched_barrier(0);
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int t = 0; t < RF; ++t) ub[v] = mfma16x16x4(lt[(S::T_FT + t) * 64], Hb[v][t], ub[v]);

    // ---------------- output ---------------------------------------------------------------
    if constexpr (DEFER) {
        if (rk) {
#pragma unroll
            for (int r = 0; r < R; ++r) rk_terms(r, adv[r], aiv[r]);
        }
    }
    if (!valid) return;  // (no barrier follows inside the tile)
    const auto r_out = make_rsrc(sp.out, state_bytes);
    const float fs = rk ? sp.f.cb * sp.f.dt : 1.0f;  // weight of (dL/du)^T lambda in the output
#pragma unroll
    for (int r = 0; r < R; ++r) {
        v4u o;
#pragma unroll
        for (int v = 0; v < 4; ++v) o[v] = __float_as_uint((rk ? xo[v][r] : 0.0f) + fs * ub[v][r]);
        __builtin_amdgcn_raw_buffer_store_b128(o, r_out, slot_off(r, unsigned(e)), 0, 0);
    }
}

// dynamic LDS of the block: the face-point records
__device__ __forceinline__ float4* vjp_records_lds() {
    extern __shared__ float4 s_vjp_rec[];
    return s_vjp_rec;
}

template <int P, bool LAT>
__device__ __forceinline__ void vjp_tile_body(const VjpParams& sp) {
    __shared__ float4 s_tab[(MfmaVjpShape<P>::TOTAL + 3) / 4];
    vjp_tile<P, LAT>(sp, xcd_tile(blockIdx.x, gridDim.x), reinterpret_cast<float*>(s_tab), vjp_records_lds());
}

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) void vjp_mfma_kernel_p3(VjpParams sp) {
    vjp_tile_body<3, false>(sp);
}
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) void vjp_mfma_kernel_p4(VjpParams sp) {
    vjp_tile_body<4, false>(sp);
}
// at most two tiles per CU: two waves per SIMD suffice (p = 3; at p = 4 this form spills)
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 8))) void vjp_mfma_lat_kernel_p3(VjpParams sp) {
    vjp_tile_body<3, true>(sp);
}

// dynamic LDS: U, P (float4 each) and the wave speed of the tile's own face points and of its
// external faces' neighbour points (at least one face)
size_t vjp_records_bytes(int p, int halo_cap) {
    const size_t nfq = size_t(p + 1);
    const size_t n = 3 * nfq * kTileEl + size_t(halo_cap > 1 ? halo_cap : 1) * nfq;
    return n * (2 * sizeof(float4) + sizeof(float));
}

}  // namespace

hipError_t launch_vjp_mfma(int p, const VjpParams& vp, hipStream_t s) {
    if (vp.K <= 0) return hipSuccess;
    if (!vp.geo32 || !vp.conn || !vp.ext || !vp.ext_off || !vp.ext_nint || !vp.halo) return hipErrorInvalidValue;
    dim3 grid((vp.K + 63) / 64), block(kBlock);
    const size_t lds = vjp_records_bytes(p, vp.halo_cap);
    // the CU count of the current device, queried once per device
    static int cus[64] = {};
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return err;
    if (dev >= 0 && dev < 64 && cus[dev] == 0) {
        int n = 0;
        err = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
        if (err != hipSuccess) return err;
        cus[dev] = n > 0 ? n : 1;
    }
    const int ncu = dev >= 0 && dev < 64 ? cus[dev] : 1;
    // p = 3 grids of at most two tiles per CU: the latency-bound form (p = 4 would spill)
    const bool lat = p == 3 && int(grid.x) <= 2 * ncu;
    switch (p) {
    case 3:
        if (lat) hipLaunchKernelGGL(vjp_mfma_lat_kernel_p3, grid, block, lds, s, vp);
        else hipLaunchKernelGGL(vjp_mfma_kernel_p3, grid, block, lds, s, vp);
        break;
    case 4: hipLaunchKernelGGL(vjp_mfma_kernel_p4, grid, block, lds, s, vp); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace dgx
