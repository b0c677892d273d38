// FP32 Euler stage kernel on the matrix cores (gfx950 v_mfma_f32_16x16x4f32), orders 3 and 4,
// volume rule exact to degree 2p.  Same operator and output modes as the per-lane stage kernel
// (dgx_stage_impl.hpp): volume residual + Rusanov faces + inverse mass + SSP-RK3 update.
//
// Reference behaviour (paths relative to the reference root):
//   volume      S/weak_forms/euler_weak_formulation.cpp:72-115
//   faces       S/weak_forms/euler_weak_formulation.cpp:117-250
//   assembly    S/solver/assembler.cpp:137-279
//   M^-1, RK3   S/solver/compressible_solver_base.cpp:54-61,106-112, I/solver/time_stepping.hpp:39-55
//
// A wave owns 16 consecutive elements (lane layout in dgx_mfma.hpp).  The four dense
// contractions of an element run as 16x16x4 products with the elements as columns, one product
// chain per conserved variable:
//   U at the volume points     = Phi_V (points x modes)   * u (modes x elements)
//   own traces at face points  = Phi_F (points x modes)   * u
//   projection                 = [PX PY]^T (modes x 2 Nq) * [F^; G^] (2 Nq x elements)
//   lift                       = LF^T (modes x 3 Nfq)     * H (3 Nfq x elements)
// The pointwise physics (contravariant fluxes, Rusanov) runs on the VALU in the slots the
// products leave in each lane.  The neighbour's trace on a face is a per-element gather with a
// per-element table (the neighbour's local face and point order), so it stays on the VALU: each
// lane group forms the partial sum over its own mode slots of the neighbour's coefficients and
// the four partials of a face point are summed through LDS.
#include "dgx_mfma_dev.hpp"

namespace dgx {
namespace {

constexpr int kTileEl = kBlock / 4;  // elements per block tile: 4 waves x 16

// LDS of one block: lane tables (and the neighbour-trace and bf16 tables), the tile's FP64
// geometry rows and the neighbour traces [face][point][element].
template <int P, bool BF16>
struct MfmaLds {
    using S = MfmaShape<P>;
    static constexpr int NTAB = BF16 ? S::TOTAL_BF : S::TOTAL;
    // p = 3: u^n of the second and third RK stages is prefetched into LDS by the tile's first
    // loads (16 B per lane and mode slot, lane-ordered), so the output does not wait a memory
    // round trip; p = 4 keeps the late register load (the LDS would cost a block per CU)
    static constexpr bool UN_LDS = P == 3;
    float tab[NTAB];
    double geo[kGeoRows][kTileEl];
    float4 tr[3 * S::NFQ * kTileEl];
    float4 un[UN_LDS ? 4 * S::R : 1][64];  // [wave * R + slot][lane]
};

// One 64-element tile starting at element eb0.  TABLES: the tile's prologue also stages the
// lane tables in LDS (a single-tile block); otherwise the caller has staged them.  The caller
// separates consecutive tiles of one block with a barrier.
template <int P, bool BF16, bool TABLES>
__device__ __forceinline__ void stage_mfma_tile(const StageParams& sp, int eb0, MfmaLds<P, BF16>& L) {
    using S = MfmaShape<P>;
    constexpr int NTAB = MfmaLds<P, BF16>::NTAB;
    constexpr int NP = S::NP, NFQ = S::NFQ, R = S::R, RQ = S::RQ, RF = S::RF;
    constexpr int EB = kTileEl;
    float* s_tab = L.tab;
    double (*s_geo)[EB] = L.geo;
    float4* s_tr = L.tr;

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int e = eb0 + 16 * w + j;  // element of this lane in the product layout
    const bool valid = e < sp.K;
    const int ec = valid ? e : sp.K - 1;  // lanes past the end compute on a valid element
    // neighbour-trace role: wave f < 3 evaluates the neighbour's trace on face f of block
    // element `lane` (one element per lane, so its gathers are as coalesced as the mesh allows)
    const int et = min(eb0 + lane, sp.K - 1);
    const unsigned uKp = unsigned(sp.Kp);
    const bool residual_mode = sp.mode == MODE_RESIDUAL;
    const float gam = float(sp.gamma);
    const float ur[4] = {float(sp.uref[0]), float(sp.uref[1]), float(sp.uref[2]), float(sp.uref[3])};
    const float* __restrict__ tm = static_cast<const float*>(sp.tabm);

    // All global traffic goes through buffer resources (the per-lane part of an address is one
    // 32-bit offset, the variable or geometry row a scalar offset) and every load is issued up
    // front in dependency order, so a wave waits for memory once.  A padding mode slot gets a
    // per-lane offset past the end of its resource, which the range check turns into a zero
    // read or a dropped write.  The tables and the block's geometry are loaded cooperatively
    // (one contiguous geometry row of 64 elements per wave instruction) and staged in LDS, so
    // no lane holds them in registers while the loads are in flight.
    constexpr int kTabPerThread = TABLES ? (NTAB + kBlock - 1) / kBlock : 0;
    float tabv[kTabPerThread > 0 ? kTabPerThread : 1];
#pragma unroll
    for (int k = 0; k < kTabPerThread; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        tabv[k] = t < NTAB ? tm[t] : 0.0f;
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
    // wave-uniform values that enter scalar buffer offsets go through readfirstlane, so the
    // loads take them from a scalar register instead of a per-lane loop
    const int wu = __builtin_amdgcn_readfirstlane(w);
    const int ft = wu < 3 ? wu : 0;
    const uint32_t ct = (__builtin_amdgcn_raw_buffer_load_b32(r_code, unsigned(et) * 4u, 0, 0) >> (4 * ft)) & 15u;
    const int nbt = int(__builtin_amdgcn_raw_buffer_load_b32(r_nbr, unsigned(et) * 4u, unsigned(ft) * 4u * uKp, 0));

    // byte offset of the four variables of mode slot (g, s) of element `el` (one 16-byte
    // access); past the end for a padding slot
    auto slot_off = [&](int s, unsigned el) -> unsigned {
        const int m = mf_mode(R, NP, g, s);
        return m >= 0 ? (unsigned(m) * uKp + el) * 16u : state_bytes;
    };
    // coefficients of this lane's mode slots (mode R g + s): evaluation operands and RK input
    float b[4][R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const v4u x = load_x4(r_uk, slot_off(s, unsigned(ec)));
#pragma unroll
        for (int v = 0; v < 4; ++v) b[v][s] = __uint_as_float(x[v]);
    }
    // u^n into LDS (p = 3; see MfmaLds): wave-uniform destination, lane L's 16 bytes at 16 L
    const bool rk_blend = sp.mode == MODE_RK && sp.ca != 0.0;
    if constexpr (MfmaLds<P, BF16>::UN_LDS) {
        if (rk_blend) {
            const auto r_un = make_rsrc(sp.un, state_bytes);
#pragma unroll
            for (int r = 0; r < R; ++r)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    r_un, (__attribute__((address_space(3))) void*)&L.un[wu * R + r][0], 16, slot_off(r, unsigned(ec)), 0,
                    0, 0);
        }
    }
    // geometry rows of the block: wave w loads rows w + 4k of the block's elements (rows past
    // the last read past the resource: zeros, no traffic)
    constexpr int kGeoPerThread = (kGeoRows * EB + kBlock - 1) / kBlock;
    double geov[kGeoPerThread];
#pragma unroll
    for (int k = 0; k < kGeoPerThread; ++k)
        geov[k] = load_f64(r_geo, unsigned(et) * 8u, unsigned(min(wu + 4 * k, kGeoRows)) * 8u * uKp);
    // neighbour coefficients for the trace role (all modes of one neighbour per lane)
    const bool t_interior = w < 3 && (ct & 3u) != kFaceBoundary;
    float cn[NP][4];
    if (t_interior) {
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const v4u x = load_x4(r_uk, (unsigned(i) * uKp + unsigned(nbt)) * 16u);
#pragma unroll
            for (int v = 0; v < 4; ++v) cn[i][v] = __uint_as_float(x[v]);
        }
    }
#pragma unroll
    for (int k = 0; k < kTabPerThread; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        if (t < NTAB) s_tab[t] = tabv[k];
    }
#pragma unroll
    for (int k = 0; k < kGeoPerThread; ++k)
        if (wu + 4 * k < kGeoRows) s_geo[wu + 4 * k][lane] = geov[k];
    __syncthreads();
    // keep the scheduler from sinking the loads above into the arithmetic below
    __builtin_amdgcn_sched_barrier(0);

    // ---------------- neighbour traces (waves 0-2) ------------------------------------------
    if (t_interior) {
        const int fn = int(ct & 3u), flip = int((ct >> 2) & 1u);
        const float* T = &s_tab[S::T_N * 64 + (fn * 2 + flip) * NP * NFQ];
        float Un[NFQ][4];
#pragma unroll
        for (int q = 0; q < NFQ; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) Un[q][v] = ur[v];
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int q = 0; q < NFQ; ++q) {
                const float t = T[i * NFQ + q];
#pragma unroll
                for (int v = 0; v < 4; ++v) Un[q][v] = fma(t, cn[i][v], Un[q][v]);
            }
#pragma unroll
        for (int q = 0; q < NFQ; ++q)
            s_tr[(w * NFQ + q) * EB + lane] = make_float4(Un[q][0], Un[q][1], Un[q][2], Un[q][3]);
    }

    const int eloc = 16 * w + j;
    double gd[kGeoRows];  // FP64 geometry of this lane's element
#pragma unroll
    for (int k = 0; k < kGeoRows; ++k) gd[k] = s_geo[k][eloc];
    // this lane's entries of the product tables (A operands)
    const float* lt = &s_tab[lane];
    float tEV[R], tEF[R], tPX[RQ], tPY[RQ], tLF[RF];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        tEV[s] = lt[(S::T_EV + s) * 64];
        tEF[s] = lt[(S::T_EF + s) * 64];
    }
#pragma unroll
    for (int t = 0; t < RQ; ++t) {
        tPX[t] = lt[(S::T_PX + t) * 64];
        tPY[t] = lt[(S::T_PY + t) * 64];
    }
#pragma unroll
    for (int t = 0; t < RF; ++t) tLF[t] = lt[(S::T_LF + t) * 64];

    const double det = gd[G_DET];
    const double rdet = 1.0 / det;
    // volume terms carry |J| only for M dU/dt (with M^-1 = M_ref^-1/|J| folded they cancel);
    // the contravariant fluxes are linear in dxi_dx, so the factor goes into its entries
    const double sv = residual_mode ? det : 1.0;
    const float a00 = float(sv * gd[G_A00]), a01 = float(sv * gd[G_A01]);
    const float a10 = float(sv * gd[G_A10]), a11 = float(sv * gd[G_A11]);
    float fnx[3], fny[3], fsf[3];  // outward normal and lift scale of each face
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        fnx[f] = float(gd[G_N0X + 2 * f]);
        fny[f] = float(gd[G_N0X + 2 * f + 1]);
        fsf[f] = residual_mode ? float(gd[G_L0 + f]) : float(gd[G_L0 + f] * rdet);
    }

    // ---------------- evaluations at the volume and face slots (matrix cores) ---------------
    v4f U[4], Uf[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        v4f c = {0.0f, 0.0f, 0.0f, 0.0f}, d = c;
#pragma unroll
        for (int s = 0; s < R; ++s) d = mfma16x16x4(tEF[s], b[v][s], d);
        if constexpr (BF16) {
            // volume evaluation from bfloat16 coefficients and basis values, FP32 accumulation
            const unsigned* eb = reinterpret_cast<const unsigned*>(&s_tab[S::T_EB + lane]);
            c = mfma16x16x16bf16(eb[0], eb[64], pack_bf16x2(b[v][0], R > 1 ? b[v][1] : 0.0f),
                                 pack_bf16x2(R > 2 ? b[v][2] : 0.0f, R > 3 ? b[v][R > 3 ? 3 : 0] : 0.0f), c);
        } else {
#pragma unroll
            for (int s = 0; s < R; ++s) c = mfma16x16x4(tEV[s], b[v][s], c);
        }
        U[v] = c;
        Uf[v] = d;
    }

    // ---------------- contravariant volume fluxes and their projection -------------------------
    // The projection contracts one register slot per product, so slot r's two products can follow
    // its fluxes directly, keeping one slot of fluxes live (the accumulation order per variable is
    // the slot order either way): p = 4 then fits four waves per SIMD instead of three; p = 3
    // times best with the products after all slots (measured).  bfloat16: one product covers
    // all four slots.
    constexpr bool kSlotProjection = !BF16 && P == 4;
    v4f acc[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    v4f Fh[4], Gh[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (r >= RQ) {
#pragma unroll
            for (int v = 0; v < 4; ++v) { Fh[v][r] = 0.0f; Gh[v][r] = 0.0f; }
            continue;
        }
        const float U0 = U[0][r] + ur[0], U1 = U[1][r] + ur[1], U2 = U[2][r] + ur[2], U3 = U[3][r] + ur[3];
        const float ri = rcp(U0);
        const float uu = U1 * ri, vv = U2 * ri;
        const float p = (gam - 1.0f) * (U3 - 0.5f * U0 * (uu * uu + vv * vv));
        // contravariant velocities along the rows of dxi_dx^T
        const float c1 = uu * a00 + vv * a10, c2 = uu * a01 + vv * a11;
        Fh[0][r] = U0 * c1;            Gh[0][r] = U0 * c2;
        Fh[1][r] = U1 * c1 + p * a00;  Gh[1][r] = U1 * c2 + p * a01;
        Fh[2][r] = U2 * c1 + p * a10;  Gh[2][r] = U2 * c2 + p * a11;
        Fh[3][r] = (U3 + p) * c1;      Gh[3][r] = (U3 + p) * c2;
        if constexpr (kSlotProjection) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                acc[v] = mfma16x16x4(tPX[r], Fh[v][r], acc[v]);
                acc[v] = mfma16x16x4(tPY[r], Gh[v][r], acc[v]);
            }
        }
    }
    if constexpr (!BF16 && !kSlotProjection) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int t = 0; t < RQ; ++t) {
                acc[v] = mfma16x16x4(tPX[t], Fh[v][t], acc[v]);
                acc[v] = mfma16x16x4(tPY[t], Gh[v][t], acc[v]);
            }
    }
    if constexpr (BF16) {
        // projection of bfloat16 fluxes on bfloat16 basis gradients, FP32 accumulation
        const unsigned* xb = reinterpret_cast<const unsigned*>(&s_tab[S::T_XB + lane]);
        const unsigned* yb = reinterpret_cast<const unsigned*>(&s_tab[S::T_YB + lane]);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            acc[v] = mfma16x16x16bf16(xb[0], xb[64], pack_bf16x2(Fh[v][0], Fh[v][1]), pack_bf16x2(Fh[v][2], Fh[v][3]), acc[v]);
            acc[v] = mfma16x16x16bf16(yb[0], yb[64], pack_bf16x2(Gh[v][0], Gh[v][1]), pack_bf16x2(Gh[v][2], Gh[v][3]), acc[v]);
        }
    }
    __syncthreads();  // neighbour traces complete

    // ---------------- faces: slot (g, f) = face f, point g (f < 3); face g, point 4 (f = 3) ---
    // Interior faces first, without branches (a boundary slot computes a value from stale LDS
    // that the pass after it replaces).
    v4f Hs[4];
    bool any_boundary = false;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const bool slot_valid = f < 3 ? g < NFQ : (NFQ == 5 && g < 3);
        const int face = f < 3 ? f : (g < 3 ? g : 0);
        const int q = f < 3 ? (g < NFQ ? g : 0) : 4;
        if (f == 3 && NFQ < 5) {
#pragma unroll
            for (int v = 0; v < 4; ++v) Hs[v][f] = 0.0f;
            continue;
        }
        const uint32_t c = (code >> (4 * face)) & 15u;
        const bool isL = (c >> 3) & 1u;
        any_boundary |= slot_valid && (c & 3u) == kFaceBoundary;
        const float nx = face == 0 ? fnx[0] : face == 1 ? fnx[1] : fnx[2];
        const float ny = face == 0 ? fny[0] : face == 1 ? fny[1] : fny[2];
        const float sf = face == 0 ? fsf[0] : face == 1 ? fsf[1] : fsf[2];
        const float4 tn = s_tr[(face * NFQ + q) * EB + eloc];
        const float Un[4] = {tn.x, tn.y, tn.z, tn.w};
        const float Uo[4] = {Uf[0][f] + ur[0], Uf[1][f] + ur[1], Uf[2][f] + ur[2], Uf[3][f] + ur[3]};
        // canonical orientation: A side = the L element, normal = the L-side outward normal
        const float sgn = isL ? 1.0f : -1.0f;
        float Aq[4], Bq[4], H[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            Aq[v] = isL ? Uo[v] : Un[v];
            Bq[v] = isL ? Un[v] : Uo[v];
        }
        rusanov2(Aq, Bq, sgn * nx, sgn * ny, gam, H);
        // a padding slot may hold a non-finite value (stale LDS): select, do not scale
        const float hs = 0.5f * sgn * sf;
#pragma unroll
        for (int v = 0; v < 4; ++v) Hs[v][f] = slot_valid ? hs * H[v] : 0.0f;
    }
    // boundary face slots: ghost state from the boundary condition (only waves that have one)
    if (__builtin_amdgcn_ballot_w64(any_boundary) != 0) {
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            if (f == 3 && NFQ < 5) continue;
            const bool slot_valid = f < 3 ? g < NFQ : (NFQ == 5 && g < 3);
            const int face = f < 3 ? f : (g < 3 ? g : 0);
            const uint32_t c = (code >> (4 * face)) & 15u;
            if (!slot_valid || (c & 3u) != kFaceBoundary) continue;
            const int nb = face == 0 ? nbf[0] : face == 1 ? nbf[1] : nbf[2];
            const float nx = face == 0 ? fnx[0] : face == 1 ? fnx[1] : fnx[2];
            const float ny = face == 0 ? fny[0] : face == 1 ? fny[1] : fny[2];
            const float sf = face == 0 ? fsf[0] : face == 1 ? fsf[1] : fsf[2];
            const float Uo[4] = {Uf[0][f] + ur[0], Uf[1][f] + ur[1], Uf[2][f] + ur[2], Uf[3][f] + ur[3]};
            float Un[4], H[4];
            const int q = f < 3 ? g : 4;  // this slot's point on the face
            face_point_flux(Uo, Un, c, bc_row(sp.bc, nb, NFQ, q), nx, ny, sf, gam, H);
#pragma unroll
            for (int v = 0; v < 4; ++v) Hs[v][f] = H[v];
        }
    }

    // lift: acc += LF^T H (LF carries -w_q/2 and, for the RK/RHS tables, M_ref^-1)
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int t = 0; t < RF; ++t) acc[v] = mfma16x16x4(tLF[t], Hs[v][t], acc[v]);

    // ---------------- output ---------------------------------------------------------------
    if constexpr (MfmaLds<P, BF16>::UN_LDS) {
        // every wave drains its u^n prefetch here, also lanes and waves past the mesh end, so no
        // LDS transfer is outstanding when a wave leaves the tile
        if (rk_blend) {
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    if (!valid) return;  // (no barrier follows inside the tile)
    const auto r_out = make_rsrc(sp.out, state_bytes);
    const float dt = float(sp.dt), ca = float(sp.ca), cb = float(sp.cb);
    // u^n for the second and third RK stages
    float unv[4][R];
    if (rk_blend) {
        if constexpr (MfmaLds<P, BF16>::UN_LDS) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float4 x = L.un[wu * R + r][lane];
                unv[0][r] = x.x; unv[1][r] = x.y; unv[2][r] = x.z; unv[3][r] = x.w;
            }
        } else {
            const auto r_un = make_rsrc(sp.un, state_bytes);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const v4u x = load_x4(r_un, slot_off(r, unsigned(e)));
#pragma unroll
                for (int v = 0; v < 4; ++v) unv[v][r] = __uint_as_float(x[v]);
            }
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

template <int P, bool BF16>
__device__ __forceinline__ void stage_mfma_body(const StageParams& sp) {
    __shared__ MfmaLds<P, BF16> L;
    stage_mfma_tile<P, BF16, true>(sp, xcd_tile(blockIdx.x, gridDim.x) * kTileEl, L);
}

// Batched instances of one order: each block walks its share of the work items (instance,
// first element, element count; counts are whole tiles except at an instance's end) with a
// fixed stride, tile by tile, the order's tables staged once.
template <int P>
__device__ __forceinline__ void batch_stage_mfma_body(const StageParams* __restrict__ params,
                                                      const BatchItem* __restrict__ items, int n_items,
                                                      const float* __restrict__ tm) {
    using S = MfmaShape<P>;
    __shared__ MfmaLds<P, false> L;
    for (int k = int(threadIdx.x); k < S::TOTAL; k += kBlock) L.tab[k] = tm[k];
    __syncthreads();
    for (int it = int(blockIdx.x); it < n_items; it += int(gridDim.x)) {
        const BatchItem item = items[it];
        const StageParams sp = params[item.instance];
        for (int e0 = item.e0; e0 < item.e0 + item.count; e0 += kTileEl) {
            stage_mfma_tile<P, false, false>(sp, e0, L);
            __syncthreads();  // the next tile overwrites the staged geometry and traces
        }
    }
}

// waves per SIMD as measured best on gfx950: p = 3 is asked for three (it gets four), p = 4
// for four
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) void stage_mfma_kernel_p3(StageParams sp) {
    stage_mfma_body<3, false>(sp);
}
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void stage_mfma_kernel_p4(StageParams sp) {
    stage_mfma_body<4, false>(sp);
}
// bfloat16 volume products (precision DGX_FP32_BF16VOL)
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) void stage_mfma_bf16_kernel_p3(StageParams sp) {
    stage_mfma_body<3, true>(sp);
}
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void stage_mfma_bf16_kernel_p4(StageParams sp) {
    stage_mfma_body<4, true>(sp);
}
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) void batch_stage_mfma_kernel_p3(
    const StageParams* params, const BatchItem* items, int n_items, const float* tab) {
    batch_stage_mfma_body<3>(params, items, n_items, tab);
}
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) void batch_stage_mfma_kernel_p4(
    const StageParams* params, const BatchItem* items, int n_items, const float* tab) {
    batch_stage_mfma_body<4>(params, items, n_items, tab);
}

}  // namespace

hipError_t launch_stage_mfma(int p, bool bf16vol, const StageParams& sp, hipStream_t s) {
    if (sp.K <= 0) return hipSuccess;
    dim3 grid((sp.K + 63) / 64), block(kBlock);
    switch (p * 2 + (bf16vol ? 1 : 0)) {
    case 6: hipLaunchKernelGGL(stage_mfma_kernel_p3, grid, block, 0, s, sp); break;
    case 7: hipLaunchKernelGGL(stage_mfma_bf16_kernel_p3, grid, block, 0, s, sp); break;
    case 8: hipLaunchKernelGGL(stage_mfma_kernel_p4, grid, block, 0, s, sp); break;
    case 9: hipLaunchKernelGGL(stage_mfma_bf16_kernel_p4, grid, block, 0, s, sp); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_batch_stage_mfma(int p, const StageParams* params, const BatchItem* items, int n_items,
                                   const void* tab, int grid, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    const float* t = static_cast<const float*>(tab);
    switch (p) {
    case 3: hipLaunchKernelGGL(batch_stage_mfma_kernel_p3, dim3(grid), dim3(kBlock), 0, s, params, items, n_items, t); break;
    case 4: hipLaunchKernelGGL(batch_stage_mfma_kernel_p4, dim3(grid), dim3(kBlock), 0, s, params, items, n_items, t); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace dgx
