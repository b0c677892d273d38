// FP32 Euler stage kernel on the matrix cores (gfx950 v_mfma_f32_16x16x4f32), orders 1 to 4,
// volume rule exact to degree 2p.  Same operator and output modes as the per-lane stage kernel
// (dgx_stage_impl.hpp): volume residual + Rusanov faces + inverse mass + SSP-RK3 update.
//
// Reference behaviour (paths relative to the reference root):
//   volume      S/weak_forms/euler_weak_formulation.cpp:72-115
//   faces       S/weak_forms/euler_weak_formulation.cpp:117-250
//   assembly    S/solver/assembler.cpp:137-279
//   M^-1, RK3   S/solver/compressible_solver_base.cpp:54-61,106-112, I/solver/time_stepping.hpp:39-55
//
// A block owns one 64-element tile of the context's element order (Morton order of the
// centroids, so a tile is a compact patch of the mesh); each wave owns 16 of its elements
// (lane layout in dgx_mfma.hpp).  The four dense contractions of an element run as 16x16x4
// products with the elements as columns, one product chain per conserved variable:
//   U at the volume points     = Phi_V (points x modes)   * u (modes x elements)
//   own traces at face points  = Phi_F (points x modes)   * u
//   projection                 = [PX PY]^T (modes x 2 Nq) * [F^; G^] (2 Nq x elements)
//   lift                       = LF^T (modes x 3 Nfq)     * H (3 Nfq x elements)
// The pointwise physics (contravariant fluxes, Rusanov) runs on the VALU in the slots the
// products leave in each lane.  Every element leaves its own traces in LDS, so a face between
// two elements of the tile reads the neighbour's trace there (both sides then see the same two
// matrix-core traces, and both copies of the face flux agree bit for bit).  Only the faces that
// leave the tile need the neighbour's trace evaluated from its coefficients, from the tile's
// list of external faces: one face point per thread, or one face per lane with each wave taking
// its own points of the face.
#include "dgx_mfma_dev.hpp"

namespace dgx {
namespace {

constexpr int kTileEl = kBlock / 4;  // elements per block tile: 4 waves x 16


// Static LDS of one block: the lane tables (with the neighbour-trace table of the external
// faces and the bf16 tables) and the tile's FP32 geometry.
template <int P, bool BF16>
struct MfmaLds {
    using S = MfmaShape<P>;
    static constexpr int NTAB = BF16 ? S::TOTAL_BF : S::TOTAL;
    static constexpr int NREC = 3 * S::NFQ * kTileEl;
    float tab[NTAB];
    float geo[kGeo32Rows][kTileEl];
};

// Trace records of the tile in dynamic LDS: state U, normal flux F(U).n along the owner's
// outward normal and wave speed |u.n| + c of every own face point ([face][point][element]),
// then those of the external faces' neighbour points ([slot][point], the reading element's
// point order): NREC + halo_cap Nfq entries per array.
struct TraceRecs {
    float4* U;
    float4* F;
    float* S;
};

// One 64-element tile (tile index `tile`, elements 64 tile ..).  The caller separates
// consecutive tiles of one block with a barrier.  LATE places the external-face traces after
// the volume work instead of before it, and issues the u^n loads of the RK blend before the face
// pass instead of at the output.  PAIR deals the external faces' points one per thread (else
// one face per lane, each wave its own points of it).
template <int P, bool BF16, bool LATE, bool PAIR>
__device__ __forceinline__ void stage_mfma_tile(const StageParams& sp, int tile, MfmaLds<P, BF16>& L,
                                                float4* __restrict__ dyn) {
    using S = MfmaShape<P>;
    constexpr int NP = S::NP, NFQ = S::NFQ, R = S::R, RQ = S::RQ, RF = S::RF;
    constexpr int EB = kTileEl;
    const int eb0 = tile * EB;

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int eloc = 16 * w + j;
    const int e = eb0 + eloc;  // element of this lane in the product layout
    const bool valid = e < sp.K;
    const int ec = valid ? e : sp.K - 1;  // lanes past the end compute on a valid element
    const unsigned uKp = unsigned(sp.Kp);
    const float gam = sp.f.gamma;
    const float ur[4] = {sp.f.uref[0], sp.f.uref[1], sp.f.uref[2], sp.f.uref[3]};
    const float* __restrict__ tm = static_cast<const float*>(sp.tabm);
    // wave-uniform values that enter scalar buffer offsets go through readfirstlane, so the
    // loads take them from a scalar register instead of a per-lane loop
    const int wu = __builtin_amdgcn_readfirstlane(w);
    const int xs = sp.ext_off[tile];     // the tile's external faces
    const int nint = sp.ext_nint[tile];  // ... of which the first nint have a neighbour
    constexpr int NREC = MfmaLds<P, BF16>::NREC;
    const int nrs = NREC + (sp.halo_cap > 1 ? sp.halo_cap : 1) * NFQ;  // records per array
    const TraceRecs H{dyn, dyn + nrs, reinterpret_cast<float*>(dyn + 2 * nrs)};

    // All global traffic goes through buffer resources (the per-lane part of an address is one
    // 32-bit offset, the variable or geometry row a scalar offset) and every load is issued up
    // front.  A padding mode slot gets a per-lane offset past the end of its resource, which
    // the range check turns into a zero read or a dropped write.  The external-face list entry
    // goes first: the neighbour-coefficient loads depend on it, and the wait for it (the memory
    // counter drains in issue order) then leaves every load issued after it in flight.
    //
    // External faces with a neighbour.  PAIR: pair t = (face hk, point hq) of the tile's
    // external faces, one per thread, dealt from wave (tile mod 4) on so that the tiles sharing a
    // CU load its four SIMDs evenly (with about half the 256 threads busy, half the issue of one
    // face per lane).  Otherwise lane k of every wave takes face k (64 at a time), each wave its
    // own points of the face.  The first 64 faces come from the fixed-stride copy of the list.
    const int npair = nint * NFQ;
    int hp = (int(threadIdx.x) + 64 * (tile & 3)) & (kBlock - 1);
    int hk = lane;
    auto face_entry = [&](int k) -> int2 {
        if (kBlock / NFQ <= 64 || k < 64) return reinterpret_cast<const int2*>(sp.halo)[tile * 64 + k];
        return k < nint ? reinterpret_cast<const int2*>(sp.ext)[xs + k] : int2{-1, 0};
    };
    const int2 hx = face_entry(PAIR ? hp / NFQ : lane);
    __builtin_amdgcn_sched_barrier(0);
    const unsigned state_bytes = unsigned(NP) * 16u * uKp;
    const auto r_uk = make_rsrc(sp.uk, state_bytes);
    const auto r_geo = make_rsrc(sp.geo32, unsigned(kGeo32Rows) * 4u * uKp);
    const auto r_conn = make_rsrc(sp.conn, 8u * uKp);
    const v2u cw = __builtin_bit_cast(v2u, __builtin_amdgcn_raw_buffer_load_b64(r_conn, unsigned(ec) * 8u, 0, 0));

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
    const bool rk_blend = sp.mode == MODE_RK && sp.f.ca != 0.0f;
    // geometry rows of the tile: wave w loads rows w + 4k (one 256-byte row per instruction;
    // rows past the last read at an offset past the resource: zeros, no traffic)
    constexpr int kGeoPerThread = (kGeo32Rows + 3) / 4;
    float geov[kGeoPerThread];
#pragma unroll
    for (int k = 0; k < kGeoPerThread; ++k) {
        const int row = wu + 4 * k;
        geov[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
            r_geo, row < kGeo32Rows ? (unsigned(row) * uKp + unsigned(eb0 + lane)) * 4u : unsigned(kGeo32Rows) * 4u * uKp,
            0, 0));
    }
    // the tables, staged in LDS by cooperative loads (so that no lane holds its table entries in
    // registers while the loads are in flight); entries past the table read at an offset past
    // the resource (zeros), so no branch stands between these loads and that wait
    constexpr int NTAB = MfmaLds<P, BF16>::NTAB;
    constexpr int kTabPerThread = (NTAB + kBlock - 1) / kBlock;
    const auto r_tab = make_rsrc(tm, unsigned(NTAB) * 4u);
    float tabv[kTabPerThread];
#pragma unroll
    for (int k = 0; k < kTabPerThread; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        tabv[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r_tab, unsigned(t) * 4u, 0, 0));
    }
    // neighbour coefficients of this thread's first external face, all modes of one neighbour
    uint32_t hcode = 0;
    float cn[NP][4];
    __builtin_amdgcn_sched_barrier(0);
    {
        const int2 x = hx;
        hcode = uint32_t(x.y);
        // no neighbour (x = -1): an offset past the resource, a zero read without traffic; the
        // mode row is a scalar offset
        const unsigned vo = x.x >= 0 ? unsigned(x.x) * 16u : state_bytes;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const v4u y = __builtin_amdgcn_raw_buffer_load_b128(r_uk, vo, unsigned(i) * 16u * uKp, 0);
#pragma unroll
            for (int v = 0; v < 4; ++v) cn[i][v] = __uint_as_float(y[v]);
        }
    }
#pragma unroll
    for (int k = 0; k < kTabPerThread; ++k) {
        const int t = int(threadIdx.x) + k * kBlock;
        if (t < NTAB) L.tab[t] = tabv[k];
    }
#pragma unroll
    for (int k = 0; k < kGeoPerThread; ++k)
        if (wu + 4 * k < kGeo32Rows) L.geo[wu + 4 * k][lane] = geov[k];
    __syncthreads();
    // keep the scheduler from sinking the loads above into the arithmetic below
    __builtin_amdgcn_sched_barrier(0);

    // ---------------- traces of the external faces ----------------------------------------
    // Run either first, or last before the barrier that publishes the records: last lets the
    // neighbour-coefficient loads arrive under the volume work, at the cost of their registers
    // staying live through it (at p = 4 that spills; p = 2 is a little slower too).
    auto external_traces = [&]() {
        if constexpr (PAIR) {
            for (;;) {
                if (hp < npair) {
                    const int fk = hp / NFQ, q = hp - fk * NFQ;  // face, point
                    const int fn = int(hcode & 3u), flip = int((hcode >> 2) & 1u);
                    const float* T = &L.tab[S::T_N * 64 + (fn * 2 + flip) * NP * NFQ];
                    // records along the neighbour's outward normal, the negated normal of the
                    // reading element's face
                    const int rel = int((hcode >> 5) & 63u), rf = int((hcode >> 3) & 3u);
                    const float hnx = -L.geo[G32_N0X + 2 * rf][rel], hny = -L.geo[G32_N0Y + 2 * rf][rel];
                    float Un[4] = {ur[0], ur[1], ur[2], ur[3]};
#pragma unroll
                    for (int i = 0; i < NP; ++i) {
                        const float t = T[i * NFQ + q];
#pragma unroll
                        for (int v = 0; v < 4; ++v) Un[v] = fma(t, cn[i][v], Un[v]);
                    }
                    float Fn[4], sw;
                    trace_record(Un, hnx, hny, gam, Fn, sw);
                    const int k = NREC + fk * NFQ + q;
                    H.U[k] = make_float4(Un[0], Un[1], Un[2], Un[3]);
                    H.F[k] = make_float4(Fn[0], Fn[1], Fn[2], Fn[3]);
                    H.S[k] = sw;
                }
                // more than 256 external-face points (tiles that are not compact patches): the
                // thread's next pair, 256 on
                hp += kBlock;
                if (hp >= npair) break;
                const int2 x = reinterpret_cast<const int2*>(sp.ext)[xs + hp / NFQ];
                hcode = uint32_t(x.y);
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    const v4u y = load_x4(r_uk, (unsigned(i) * uKp + unsigned(x.x)) * 16u);
#pragma unroll
                    for (int v = 0; v < 4; ++v) cn[i][v] = __uint_as_float(y[v]);
                }
            }
        } else {
            for (;;) {
                if (hk < nint) {
                    const int fn = int(hcode & 3u), flip = int((hcode >> 2) & 1u);
                    const float* T = &L.tab[S::T_N * 64 + (fn * 2 + flip) * NP * NFQ];
                    // records along the neighbour's outward normal, the negated normal of the
                    // reading element's face
                    const int rel = int((hcode >> 5) & 63u), rf = int((hcode >> 3) & 3u);
                    const float hnx = -L.geo[G32_N0X + 2 * rf][rel], hny = -L.geo[G32_N0Y + 2 * rf][rel];
#pragma unroll
                    for (int q0 = 0; q0 < NFQ; q0 += 4) {
                        const int q = q0 + wu;
                        if (q >= NFQ) continue;
                        float Un[4] = {ur[0], ur[1], ur[2], ur[3]};
#pragma unroll
                        for (int i = 0; i < NP; ++i) {
                            const float t = T[i * NFQ + q];
#pragma unroll
                            for (int v = 0; v < 4; ++v) Un[v] = fma(t, cn[i][v], Un[v]);
                        }
                        float Fn[4], sw;
                        trace_record(Un, hnx, hny, gam, Fn, sw);
                        const int k = NREC + hk * NFQ + q;
                        H.U[k] = make_float4(Un[0], Un[1], Un[2], Un[3]);
                        H.F[k] = make_float4(Fn[0], Fn[1], Fn[2], Fn[3]);
                        H.S[k] = sw;
                    }
                }
                // more than 64 such faces (tiles that are not compact patches): the next 64
                hk += 64;
                if (hk - lane >= nint) break;
                if (hk < nint) {
                    const int2 x = reinterpret_cast<const int2*>(sp.ext)[xs + hk];
                    hcode = uint32_t(x.y);
#pragma unroll
                    for (int i = 0; i < NP; ++i) {
                        const v4u y = load_x4(r_uk, (unsigned(i) * uKp + unsigned(x.x)) * 16u);
#pragma unroll
                        for (int v = 0; v < 4; ++v) cn[i][v] = __uint_as_float(y[v]);
                    }
                }
            }
        }
    };
    if constexpr (!LATE) external_traces();

    // this lane's entries of the product tables (A operands)
    const float* lt = &L.tab[lane];
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
    // this lane's element geometry (FP32 roundings of the FP64 geometry)
    const float a00 = L.geo[G32_A00][eloc], a01 = L.geo[G32_A01][eloc];
    const float a10 = L.geo[G32_A10][eloc], a11 = L.geo[G32_A11][eloc];
    float fnx[3], fny[3], fsf[3];  // outward normal and lift scale of each face
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        fnx[f] = L.geo[G32_N0X + 2 * f][eloc];
        fny[f] = L.geo[G32_N0Y + 2 * f][eloc];
        fsf[f] = L.geo[G32_S0 + f][eloc];
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
            const unsigned* eb = reinterpret_cast<const unsigned*>(&L.tab[S::T_EB + lane]);
            c = mfma16x16x16bf16(eb[0], eb[64], pack_bf16x2(b[v][0], R > 1 ? b[v][1] : 0.0f),
                                 pack_bf16x2(R > 2 ? b[v][2] : 0.0f, R > 3 ? b[v][R > 3 ? 3 : 0] : 0.0f), c);
        } else {
#pragma unroll
            for (int s = 0; s < R; ++s) c = mfma16x16x4(tEV[s], b[v][s], c);
        }
        U[v] = c;
        Uf[v] = d;
    }
    // trace records of the own face points into LDS, for this element's faces and its tile
    // neighbours: slot (g, f) = face f, point g (f < 3); face g, point 4 (f = 3, p = 4)
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        if (f == 3 && NFQ < 5) continue;
        const bool slot_valid = f < 3 ? g < NFQ : g < 3;
        const int face = f < 3 ? f : (g < 3 ? g : 0), q = f < 3 ? (g < NFQ ? g : 0) : 4;
        const float Uo[4] = {Uf[0][f] + ur[0], Uf[1][f] + ur[1], Uf[2][f] + ur[2], Uf[3][f] + ur[3]};
        const float nx = face == 0 ? fnx[0] : face == 1 ? fnx[1] : fnx[2];
        const float ny = face == 0 ? fny[0] : face == 1 ? fny[1] : fny[2];
        float Fn[4], sw;
        trace_record(Uo, nx, ny, gam, Fn, sw);
        if (slot_valid) {
            const int k = (face * NFQ + q) * EB + eloc;
            H.U[k] = make_float4(Uo[0], Uo[1], Uo[2], Uo[3]);
            H.F[k] = make_float4(Fn[0], Fn[1], Fn[2], Fn[3]);
            H.S[k] = sw;
        }
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
        const unsigned* xb = reinterpret_cast<const unsigned*>(&L.tab[S::T_XB + lane]);
        const unsigned* yb = reinterpret_cast<const unsigned*>(&L.tab[S::T_YB + lane]);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            acc[v] = mfma16x16x16bf16(xb[0], xb[64], pack_bf16x2(Fh[v][0], Fh[v][1]), pack_bf16x2(Fh[v][2], Fh[v][3]), acc[v]);
            acc[v] = mfma16x16x16bf16(yb[0], yb[64], pack_bf16x2(Gh[v][0], Gh[v][1]), pack_bf16x2(Gh[v][2], Gh[v][3]), acc[v]);
        }
    }
    if constexpr (LATE) external_traces();
    __syncthreads();  // own traces and external-face traces complete
    // LATE: u^n for the second and third RK stages, issued here so that the face pass covers its
    // latency (a resource of size zero when there is no blend: zero reads without traffic)
    const auto r_un = make_rsrc(rk_blend ? sp.un : sp.uk, rk_blend ? state_bytes : 0u);
    v4u unx[LATE ? R : 1];
    if constexpr (LATE) {
#pragma unroll
        for (int r = 0; r < R; ++r) unx[r] = load_x4(r_un, slot_off(r, unsigned(ec)));
        __builtin_amdgcn_sched_barrier(0);
    }

    // ---------------- faces: slot (g, f) = face f, point g (f < 3); face g, point 4 (f = 3) ---
    // From the two trace records of a face point, own (o) and neighbour (n), each with its normal
    // flux along its own outward normal (exact negatives of each other): twice the Rusanov flux
    // in this element's orientation is X = (Fn_o - Fn_n) - s (U_n - U_o), s = max(s_L, s_R) with
    // the canonical side first, and the two elements of a face form exactly opposite X.
    // Interior faces first, without branches (a boundary slot computes a value from its own
    // record that the pass after it replaces).
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
        const uint32_t c = conn_field(cw[0], cw[1], face);
        const bool bnd = (c & kConnBoundary) != 0, isL = (c >> 13) & 1u;
        any_boundary |= slot_valid && bnd;
        // the neighbour's record at this point (a boundary face reads its own record as a
        // placeholder)
        const int ko = (face * NFQ + q) * EB + eloc;
        const uint32_t stepc = (c >> 11) & 3u;
        const int step = stepc == 0u ? EB : stepc == 1u ? -EB : 1;
        const int kn = bnd ? ko : int(c & 2047u) + q * step;
        const float4 uo = H.U[ko], fo = H.F[ko];
        const float so = H.S[ko];
        const float4 un = H.U[kn], fm = H.F[kn];
        const float sn = H.S[kn];
        const float sw = ref_max(isL ? so : sn, isL ? sn : so);
        const float hs = 0.5f * (face == 0 ? fsf[0] : face == 1 ? fsf[1] : fsf[2]);
        const float X0 = (fo.x - fm.x) - sw * (un.x - uo.x);
        const float X1 = (fo.y - fm.y) - sw * (un.y - uo.y);
        const float X2 = (fo.z - fm.z) - sw * (un.z - uo.z);
        const float X3 = (fo.w - fm.w) - sw * (un.w - uo.w);
        // a padding slot may hold a non-finite value: select, do not scale
        Hs[0][f] = slot_valid ? hs * X0 : 0.0f;
        Hs[1][f] = slot_valid ? hs * X1 : 0.0f;
        Hs[2][f] = slot_valid ? hs * X2 : 0.0f;
        Hs[3][f] = slot_valid ? hs * X3 : 0.0f;
    }
    // boundary face slots: ghost state from the boundary condition (only waves that have one)
    if (__builtin_amdgcn_ballot_w64(any_boundary) != 0) {
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            if (f == 3 && NFQ < 5) continue;
            const bool slot_valid = f < 3 ? g < NFQ : (NFQ == 5 && g < 3);
            const int face = f < 3 ? f : (g < 3 ? g : 0);
            const uint32_t c = conn_field(cw[0], cw[1], face);
            if (!slot_valid || !(c & kConnBoundary)) continue;
            const int bf = reinterpret_cast<const int2*>(sp.ext)[xs + int(c & 2047u)].x;
            const float nx = face == 0 ? fnx[0] : face == 1 ? fnx[1] : fnx[2];
            const float ny = face == 0 ? fny[0] : face == 1 ? fny[1] : fny[2];
            const float sf = face == 0 ? fsf[0] : face == 1 ? fsf[1] : fsf[2];
            const int q = f < 3 ? g : 4;  // this slot's point on the face
            const float4 uo = H.U[(face * NFQ + q) * EB + eloc];
            const float Uo[4] = {uo.x, uo.y, uo.z, uo.w};
            float Un[4], Hb[4];
            face_point_flux(Uo, Un, kFaceBoundary | 8u, bc_row(sp.bc, bf, NFQ, q), nx, ny, sf, gam, Hb);
#pragma unroll
            for (int v = 0; v < 4; ++v) Hs[v][f] = Hb[v];
        }
    }

    // lift: acc += LF^T H (LF carries -w_q/2 and, for the RK/RHS tables, M_ref^-1)
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int t = 0; t < RF; ++t) acc[v] = mfma16x16x4(tLF[t], Hs[v][t], acc[v]);

    // ---------------- output ---------------------------------------------------------------
    if (!valid) return;  // (no barrier follows inside the tile)
    const auto r_out = make_rsrc(sp.out, state_bytes);
    const float dt = sp.f.dt, ca = sp.f.ca, cb = sp.f.cb;
    auto store = [&](auto&& value) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            v4u x;
#pragma unroll
            for (int v = 0; v < 4; ++v) x[v] = __float_as_uint(value(v, r));
            // padding slots address past the end and are dropped
            __builtin_amdgcn_raw_buffer_store_b128(x, r_out, slot_off(r, unsigned(e)), 0, 0);
        }
    };
    if (sp.mode != MODE_RK) {
        store([&](int v, int r) { return acc[v][r]; });
    } else if (!rk_blend) {
        store([&](int v, int r) { return b[v][r] + dt * acc[v][r]; });
    } else {
        // u^n for the second and third RK stages
        float unv[4][R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const v4u x = LATE ? unx[LATE ? r : 0] : load_x4(r_un, slot_off(r, unsigned(e)));
#pragma unroll
            for (int v = 0; v < 4; ++v) unv[v][r] = __uint_as_float(x[v]);
        }
        store([&](int v, int r) { return ca * unv[v][r] + cb * (b[v][r] + dt * acc[v][r]); });
    }
}

// dynamic LDS of the block: the external-face traces
__device__ __forceinline__ float4* halo_lds() {
    extern __shared__ float4 s_halo[];
    return s_halo;
}

// single contexts: the latency order at p = 3 (measured 4% faster there, slower at p = 2, 4);
// the external faces' points dealt one per thread except at p = 3 in FP32 (there its registers
// or the uneven waves before the records barrier make it 4% slower; 1-3.5% faster elsewhere)
template <int P, bool BF16>
__device__ __forceinline__ void stage_mfma_body(const StageParams& sp) {
    __shared__ MfmaLds<P, BF16> L;
    stage_mfma_tile<P, BF16, P == 3, BF16 || P != 3>(sp, xcd_tile(blockIdx.x, gridDim.x), L, halo_lds());
}

// Batched instances of one order: each block walks its share of the work items (instance,
// first element, element count; items start on a tile and cover whole tiles except at an
// instance's end) with a fixed stride, tile by tile.
template <int P>
__device__ __forceinline__ void batch_stage_mfma_body(const StageParams* __restrict__ params,
                                                      const BatchItem* __restrict__ items, int n_items) {
    __shared__ MfmaLds<P, false> L;
    for (int it = int(blockIdx.x); it < n_items; it += int(gridDim.x)) {
        const BatchItem item = items[it];
        const StageParams& sp = params[item.instance];
        for (int e0 = item.e0; e0 < item.e0 + item.count; e0 += kTileEl) {
            stage_mfma_tile<P, false, false, true>(sp, e0 / kTileEl, L, halo_lds());
            __syncthreads();  // the next tile overwrites the staged geometry and traces
        }
    }
}

// waves per SIMD as measured best on gfx950: p = 3 is asked for three (it gets four), p = 4
// for four; p = 1, 2 for four; the batched p = 3 kernel for four (three at 129 VGPRs otherwise)
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void stage_mfma_kernel_p1(StageParams sp) {
    stage_mfma_body<1, false>(sp);
}
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void stage_mfma_kernel_p2(StageParams sp) {
    stage_mfma_body<2, false>(sp);
}
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
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void batch_stage_mfma_kernel_p1(
    const StageParams* params, const BatchItem* items, int n_items) {
    batch_stage_mfma_body<1>(params, items, n_items);
}
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void batch_stage_mfma_kernel_p2(
    const StageParams* params, const BatchItem* items, int n_items) {
    batch_stage_mfma_body<2>(params, items, n_items);
}
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void batch_stage_mfma_kernel_p3(
    const StageParams* params, const BatchItem* items, int n_items) {
    batch_stage_mfma_body<3>(params, items, n_items);
}
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) void batch_stage_mfma_kernel_p4(
    const StageParams* params, const BatchItem* items, int n_items) {
    batch_stage_mfma_body<4>(params, items, n_items);
}

// dynamic LDS of the trace records: the tile's own face points and at least one external
// face, U, F(U).n, wave speed
size_t halo_bytes(int p, int halo_cap) {
    const size_t n = size_t(3 * (p + 1) * kTileEl) + size_t(halo_cap > 1 ? halo_cap : 1) * size_t(p + 1);
    return n * (2 * sizeof(float4) + sizeof(float));
}

}  // namespace

hipError_t launch_stage_mfma(int p, bool bf16vol, const StageParams& sp, hipStream_t s) {
    if (sp.K <= 0) return hipSuccess;
    if (!sp.geo32 || !sp.conn || !sp.ext || !sp.ext_off || !sp.ext_nint) return hipErrorInvalidValue;
    dim3 grid((sp.K + 63) / 64), block(kBlock);
    const size_t lds = halo_bytes(p, sp.halo_cap);
    switch (p * 2 + (bf16vol ? 1 : 0)) {
    case 2: hipLaunchKernelGGL(stage_mfma_kernel_p1, grid, block, lds, s, sp); break;
    case 4: hipLaunchKernelGGL(stage_mfma_kernel_p2, grid, block, lds, s, sp); break;
    case 6: hipLaunchKernelGGL(stage_mfma_kernel_p3, grid, block, lds, s, sp); break;
    case 7: hipLaunchKernelGGL(stage_mfma_bf16_kernel_p3, grid, block, lds, s, sp); break;
    case 8: hipLaunchKernelGGL(stage_mfma_kernel_p4, grid, block, lds, s, sp); break;
    case 9: hipLaunchKernelGGL(stage_mfma_bf16_kernel_p4, grid, block, lds, s, sp); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_batch_stage_mfma(int p, const StageParams* params, const BatchItem* items, int n_items, int halo_cap,
                                   int grid, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    const size_t lds = halo_bytes(p, halo_cap);
    switch (p) {
    case 1: hipLaunchKernelGGL(batch_stage_mfma_kernel_p1, dim3(grid), dim3(kBlock), lds, s, params, items, n_items); break;
    case 2: hipLaunchKernelGGL(batch_stage_mfma_kernel_p2, dim3(grid), dim3(kBlock), lds, s, params, items, n_items); break;
    case 3: hipLaunchKernelGGL(batch_stage_mfma_kernel_p3, dim3(grid), dim3(kBlock), lds, s, params, items, n_items); break;
    case 4: hipLaunchKernelGGL(batch_stage_mfma_kernel_p4, dim3(grid), dim3(kBlock), lds, s, params, items, n_items); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace dgx
