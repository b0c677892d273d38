// Lane layout of the MFMA stage kernel (dgx_stage_mfma.hip) and of its lane-major tables,
// shared by the kernel and the host-side table builder.
//
// One wave processes 16 elements with v_mfma_f32_16x16x4f32 (D = A B + C, A 16x4, B 4x16):
//   A[i][k] comes from lane i + 16k, B[k][j] from lane j + 16k, and lane j + 16g holds
//   D[4g + r][j] in its register r.
// Column j of every product is element j of the tile and lane group g = lane / 16 plays the
// role of k (operands) and of the row block (results), so a lane (g, j) only ever holds data
// of element j.  The three index sets are dealt over the (g, r) "slots" of a lane:
//   modes          slot (g, r) -> mode R g + r        (r < R = ceil(Np / 4))
//   volume points  slot (g, r) -> point RQ g + r      (r < RQ = ceil(Nq / 4))
//   face points    slot (g, r), r < 3 -> face r, point g; r = 3 -> face g, point 4 (p = 4)
// With the contraction index k dealt the same way, the result registers of one product are
// the operand registers of the next: the coefficients loaded for the volume evaluation are the
// ones the RK update needs, the fluxes at the volume points are the projection's operands and
// the face fluxes are the lift's.
#pragma once
#include "dgx_internal.hpp"

namespace dgx {

template <int P>
struct MfmaShape {
    static constexpr int NP = np_of(P);
    static constexpr int NQ = nq_of(P, 0);
    static constexpr int NFQ = P + 1;
    static constexpr int R = (NP + 3) / 4;
    static constexpr int RQ = (NQ + 3) / 4;
    static constexpr int RF = NFQ > 4 ? 4 : 3;
    static_assert(R <= 4 && RQ <= 4 && NFQ <= 5, "one 16x16 tile per operand");
    // lane tables, 64 floats each: EV[R] EF[R] PX[RQ] PY[RQ] LF[RF], then the neighbour-trace
    // table NT[fn*2 + flip][i][q]: mode i of the neighbour at point q of the reading element's
    // face (the neighbour's local face fn, its points reversed when flip is set)
    static constexpr int T_EV = 0;
    static constexpr int T_EF = T_EV + R;
    static constexpr int T_PX = T_EF + R;
    static constexpr int T_PY = T_PX + RQ;
    static constexpr int T_LF = T_PY + RQ;
    static constexpr int T_N = T_LF + RF;
    static constexpr int NT = 6 * NP * NFQ;
    static constexpr int TOTAL = T_N * 64 + NT;
    // bfloat16 operands of the volume products (v_mfma_f32_16x16x16_bf16: lane i + 16 g holds
    // A[i][4g .. 4g+3]), two 32-bit words of packed pairs per lane and table, placed after the
    // FP32 tables: EB (Phi_V, k = mode slot (g, t)), XB / YB (PX / PY, rows = mode slots,
    // k = volume slot (g, t))
    static constexpr int T_EB = TOTAL;
    static constexpr int T_XB = T_EB + 2 * 64;
    static constexpr int T_YB = T_XB + 2 * 64;
    static constexpr int TOTAL_BF = T_YB + 2 * 64;
};

// Lane tables of the matrix-core stage adjoint (dgx_vjp_mfma.hip), 64 floats each:
//   EV[R]  Phi_V  (rows: volume slots, k: modes)     U at the volume points
//   XV[R]  PX     (rows: volume slots, k: modes)     lambda projected for the x-flux adjoint
//   YV[R]  PY     (rows: volume slots, k: modes)     lambda projected for the y-flux adjoint
//   VT[RQ] Phi_V  (rows: modes, k: volume slots)     transposed evaluation
//   EF[R]  Phi_F  (rows: face slots, k: modes)       own traces
//   LE[R]  LF     (rows: face slots, k: modes)       own lift weights of lambda
//   FT[RF] Phi_F  (rows: modes, k: face slots)       transposed trace evaluation
// then the external-face tables: per point q of face f (index f Nfq + q) HV float4, entry m
// holding (Phi_F, LF) of modes 2m and 2m + 1 at the point (zeros past the last mode).  Padding
// rows of EV / EF repeat point 0 (finite states); every other padding entry is zero.
template <int P>
struct MfmaVjpShape {
    using F = MfmaShape<P>;
    static constexpr int R = F::R, RQ = F::RQ, RF = F::RF;
    static constexpr int T_EV = 0;
    static constexpr int T_XV = T_EV + R;
    static constexpr int T_YV = T_XV + R;
    static constexpr int T_VT = T_YV + R;
    static constexpr int T_EF = T_VT + RQ;
    static constexpr int T_LE = T_EF + R;
    static constexpr int T_FT = T_LE + R;
    static constexpr int T_N = T_FT + RF;
    static constexpr int NP = F::NP, NFQ = F::NFQ;
    static constexpr int HV = (NP + 1) / 2;        // float4 per external-face table point
    static constexpr int TOTAL = T_N * 64 + 3 * NFQ * HV * 4;
};

// Tables of the matrix-core Navier-Stokes tile kernel (dgx_stage_mfma_ns.hip).
//
// Face gradient.  Component c of the gradient at face point q is
//   g_c(q) = sum_i sum_k dxi_dx[d][k] T_ck(q, i) u_i,
// d = c under the plain read (T_0k = T_1k = dphi_i/dxi_k at q); under the reference's read, a
// column-major view of the interleaved derivative storage, T_ck(q, i) is the derivative of
// another mode at another point and d = e(c, q) xor (i mod 2) with the row parity
// e(c, q) = (q Np + c Nfq Np) mod 2 (the builder checks both forms).
//
// Face traces (one buffer): the lane tables EF[R] (Phi_F) and FG[4][R] (T_ck, slot c*2 + k;
// rows: face slots, k: modes), 64 floats each; then the external-face tables, per point q of
// face f (index f Nfq + q) HP float4:
//   [i]       (T_0i0, T_0i1, T_1i0, T_1i1) of mode i at the point
//   [NP ..]   Phi_F of the point, modes 0 .. NP-1 (padded with zeros to whole float4)
// Volume (one buffer per mode): EV[R], GX[R], GY[R] (Phi_V and its reference derivatives; rows:
// volume slots, k: modes), PX[RQ], PY[RQ] (rows: modes, k: volume slots), LF[RF] (rows: modes,
// k: face slots).
template <int P>
struct MfmaNsShape {
    using F = MfmaShape<P>;
    static constexpr int R = F::R, RQ = F::RQ, RF = F::RF, NP = F::NP, NFQ = F::NFQ;
    static constexpr int T_EF = 0;
    static constexpr int T_FG = T_EF + R;
    static constexpr int TOTAL_T = (T_FG + 4 * R) * 64;
    static constexpr int HP = NP + (NP + 3) / 4;     // float4 per external-face table point
    static constexpr int NHT = 3 * NFQ * HP * 4;     // floats of the external-face tables
    static constexpr int TOTAL_TH = TOTAL_T + NHT;   // floats of the face-trace buffer
    static constexpr int M_EV = 0;
    static constexpr int M_GX = M_EV + R;
    static constexpr int M_GY = M_GX + R;
    static constexpr int M_PX = M_GY + R;
    static constexpr int M_PY = M_PX + RQ;
    static constexpr int M_LF = M_PY + RQ;
    static constexpr int TOTAL_M = (M_LF + RF) * 64;
};

// slot -> mode / volume point / face point (f * nfq + q); -1 for a padding slot
__host__ __device__ constexpr int mf_mode(int R, int np, int g, int r) {
    return (r < R && R * g + r < np) ? R * g + r : -1;
}
__host__ __device__ constexpr int mf_vpoint(int RQ, int nq, int g, int r) {
    return (r < RQ && RQ * g + r < nq) ? RQ * g + r : -1;
}
__host__ __device__ constexpr int mf_fpoint(int nfq, int g, int r) {
    return r < 3 ? (g < nfq ? r * nfq + g : -1) : ((nfq == 5 && g < 3) ? g * nfq + 4 : -1);
}

// Orders with a matrix-core FP32 Euler stage (tile) kernel, volume rule exact to degree 2p.
constexpr bool mfma_euler_order(int p) { return p >= 1 && p <= 4; }
// Orders with the matrix-core Navier-Stokes stage, the stage adjoint and the bfloat16 volume
// variant.
constexpr bool mfma_order(int p) { return p == 3 || p == 4; }
int mfma_table_size(int p);  // floats
// Build the lane-major tables of order p from the (L or R) tables in TabLayout order.
// With bf16 set, the bfloat16 volume tables follow (TOTAL_BF floats in all).
void build_mfma_tables(int p, const std::vector<double>& tab, std::vector<float>& out, bool bf16 = false);
void build_mfma_vjp_tables(int p, const std::vector<double>& tab, std::vector<float>& out);
// Navier-Stokes: face-trace and external-face tables (trace = true) or volume tables from the
// L / R tables.
void build_mfma_ns_tables(int p, const std::vector<double>& tab, bool trace, bool quirk, std::vector<float>& out);

}  // namespace dgx
