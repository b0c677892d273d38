// Private definitions shared by the host API (dgx_api.cpp) and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dgx.h"

namespace dgx {

constexpr int kBlock = 256;  // threads per block of the element kernels (4 wave64)

// Blocks are dealt round-robin over the GPU's 8 XCDs (block b to XCD b mod 8), each with its own
// L2.  Renumbering block b to xcd_tile(b, n) gives every XCD one contiguous range of the n
// blocks (the first n mod 8 XCDs one more), so that the mesh neighbours of a block's elements
// -- the neighbouring tiles, the mesh rows above and below -- are mostly read through the same
// L2.  A bijection of [0, n) for every n; it changes locality only.
__host__ __device__ constexpr int xcd_tile(int b, int n) {
    const int q = n / 8, r = n % 8, x = b % 8, k = b / 8;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Volume points of the Dunavant rule exact to degree 2p (+1 for the 2P1 variant).
constexpr int nq_of(int p, int rule) {
    return rule == 0 ? (p == 1 ? 3 : p == 2 ? 6 : p == 3 ? 12 : 16)
                     : (p == 1 ? 4 : p == 2 ? 7 : p == 3 ? 13 : 19);
}
constexpr int np_of(int p) { return (p + 1) * (p + 2) / 2; }

// Device state layout: coefficient i of variable v of element e at u[(i * Kp + e) * 4 + v] --
// mode-major rows of Kp elements, the four conserved variables of an element contiguous (one
// 16-byte access per element and mode in FP32).
__host__ __device__ constexpr int sidx(int i, int v, int e, int Kp) { return (i * Kp + e) * 4 + v; }

// Geometry rows of the per-element FP64 array geo[kGeoRows][Kp].
enum GeoRow : int {
    G_A00 = 0, G_A01, G_A10, G_A11,  // dxi_dx = inverse of J^T (row-major)
    G_DET,                            // |det J|
    G_N0X, G_N0Y, G_N1X, G_N1Y, G_N2X, G_N2Y,  // outward unit normals of faces 0..2
    G_L0, G_L1, G_L2,                 // face lengths
    kGeoRows
};

// Per-element face code: 4 bits per local face f at bit 4f:
//   bits 0-1 neighbour local face (3 = boundary face), bit 2 reversed point order,
//   bit 3 this element holds the canonical (L) orientation of the face.
constexpr uint32_t kFaceBoundary = 3u;

// Boundary data per boundary-face point.  A boundary face's entry in the neighbour array is its
// running index bf; face point q of it owns the row bc + (bf * nfq + q) * kBcRow:
//   [0]    condition type (dgx_bc_type; -1: no condition, the face contributes nothing)
//   [1..4] the Euler ghost data (FAR_FIELD conserved, NO_SLIP primitive) evaluated at the
//          point the reference's Euler boundary face reads (its transposed read of the face
//          points, euler_weak_formulation.cpp:188-194, or the point itself)
//   [5..8] the same condition at the point itself, for the viscous boundary flux
//          (navier_stokes_weak_formulation.cpp:244-264)
constexpr int kBcRow = 9;
__host__ __device__ inline const double* bc_row(const double* bc, int bf, int nfq, int q) {
    return bc + (size_t(bf) * nfq + q) * kBcRow;
}

// Table layout (Real), offsets in elements.  "L" tables carry the reference-element mass
// inverse folded in; "R" tables produce M dU/dt.
struct TabLayout {
    int np, nq, nfq;
    int phiV, PX, PY, phiF, LF, DXV, DYV, FG, mass, total;
    TabLayout(int p, int rule) {
        np = np_of(p);
        nq = nq_of(p, rule);
        nfq = p + 1;
        int o = 0;
        phiV = o; o += nq * np;
        PX = o; o += nq * np;
        PY = o; o += nq * np;
        phiF = o; o += 3 * nfq * np;
        LF = o; o += 3 * nfq * np;
        DXV = o; o += nq * np;
        DYV = o; o += nq * np;
        FG = o; o += 3 * nfq * 8 * np;  // face gradients [3][nfq][slot (c,d,k)][np]
        mass = o; o += np;
        total = o;
    }
};

// Quadrilateral elements of order p (tensor Legendre basis, (p+1)^2 tensor GL points):
// table offsets (state precision, plus an FP64 copy) and FP64 geometry rows [row][Kp].
//   tables: phiV [nq][np]; GXV, GYV [nq][np] reference gradients; phiF [4][nfq][np];
//           LF [4][nfq][np] = -w_q/2 phi; GF [4][nfq][np][2] reference gradients at face points
//   geometry: per volume point q rows gV+5q..: dxi/dx (A00 A01 A10 A11), w_q |J_q|;
//             per face f rows gF+3f..: outward normal (2), length (the canonical side's on both
//             elements of an interior face, negated normal on the other side);
//             per face point (f,q) rows gFA+4(f nfq+q)..: dxi/dx
struct QuadDims {
    int np, nq, nfq;
    int phiV, GXV, GYV, phiF, LF, GF, T1, D1, X1, W1, IM1, FX, FY, LX, total;
    int gV, gF, gFA, gX, rows;
    __host__ __device__ constexpr explicit QuadDims(int p)
        : np((p + 1) * (p + 1)), nq((p + 1) * (p + 1)), nfq(p + 1), phiV(0), GXV(nq * np), GYV(2 * nq * np),
          phiF(3 * nq * np), LF(3 * nq * np + 4 * nfq * np), GF(3 * nq * np + 8 * nfq * np),
          T1(3 * nq * np + 16 * nfq * np), D1(T1 + nfq * nfq), X1(D1 + nfq * nfq), W1(X1 + nfq), IM1(W1 + nfq),
          FX(IM1 + nfq), FY(FX + 4 * nfq * nfq), LX(FY + 4 * nfq * nfq), total(LX + 4 * nfq * nfq), gV(0),
          gF(5 * nq), gFA(5 * nq + 12), gX(5 * nq + 12 + 16 * nfq), rows(5 * nq + 12 + 16 * nfq + 8) {}
};
// Factors of the tensor basis for the four-lane quad kernel (offsets above, R = p + 1 GL nodes
// s_q with weights w_q; P_a the degree-a Legendre polynomial, mode (a, b) = index b R + a):
//   T1[a R + q] = P_a(s_q), D1[a R + q] = P_a'(s_q), X1[q] = s_q, W1[q] = w_q,
//   IM1[a] = 1 / sum_q w_q P_a(s_q)^2,
//   FX[(f R + q) R + a] = P_a(xi) and FY[(f R + q) R + b] = P_b(eta) at face point q of face f,
//   LX[(f R + q) R + a] = -w_q / 2 P_a(xi) there (the lift's xi factor).
// Geometry rows gX .. gX + 7: the element's vertices x0, y0, .., x3, y3.

// Quad element code: bits [5f, 5f+5) describe face f (neighbour's local face in 2 bits, reversed
// point order, canonical (L) side, boundary); bit 20 marks a parallelogram (constant Jacobian,
// diagonal mass matrix in the Legendre basis).
__host__ __device__ inline uint32_t quad_code(uint32_t code, int f) { return (code >> (5 * f)) & 31u; }
constexpr uint32_t kQuadAffineBit = 1u << 20;

// FP32 per-element geometry of the matrix-core tile kernels, rows [kGeo32NsRows][Kp]: the FP32
// roundings of the FP64 geometry as the kernels consume them, formed in FP64 on the device
// (dgx_geometry.hip): dxi_dx (times |det J| for M dU/dt), the three outward normals, and the
// three lift scales len_f / |det J| (len_f for M dU/dt) -- the Euler kernel's kGeo32Rows rows --
// then, for the Navier-Stokes kernel, dxi_dx itself (the gradients) and the face lengths (the
// penalty).
enum Geo32Row : int { G32_A00 = 0, G32_A01, G32_A10, G32_A11, G32_N0X, G32_N0Y, G32_N1X, G32_N1Y, G32_N2X, G32_N2Y,
                      G32_S0, G32_S1, G32_S2, kGeo32Rows,
                      G32_G00 = kGeo32Rows, G32_G01, G32_G10, G32_G11, G32_L0, G32_L1, G32_L2, kGeo32NsRows };

// Face connections of the tile kernel, per element two 32-bit words, 16 bits per face f (word
// f >> 1, shifted by 16 (f & 1)), addressing the tile's trace records: the records of the
// tile's own face points at index (face Nfq + point) 64 + element, those of the external faces'
// neighbour points after them at 3 Nfq 64 + slot Nfq + point (in the reading element's order).
//   bits 0-10   index of the neighbour's record at point 0 of the face (boundary face: its slot
//               in the tile's external-face list)
//   bits 11-12  step of that index per point: 0 = +64, 1 = -64 (reversed order), 2 = +1
//   bit 13      this element holds the canonical (L) orientation
//   bit 14      boundary face
// External faces of tile t are entries ext_off[t] .. ext_off[t+1]-1 of a list of int2 (x: the
// neighbour element, or the boundary face's running index; y: neighbour's local face | reversed
// order << 2 | the reading element's local face << 3 | its index in the tile << 5 | the reading
// element holds the canonical side << 11), its ext_nint[t] faces with a neighbour first.
constexpr uint32_t kConnBoundary = 1u << 14;
__host__ __device__ constexpr uint32_t conn_field(uint32_t w0, uint32_t w1, int f) {
    return ((f < 2 ? w0 : w1) >> (16 * (f & 1))) & 0xffffu;
}

// Launch parameters of one fused stage: out = ca*un + cb*(uk + dt*L(uk)) (mode RK),
// out = L(uk) (mode RHS) or out = M dU/dt (mode RESIDUAL).
enum StageMode : int { MODE_RK = 0, MODE_RHS = 1, MODE_RESIDUAL = 2 };

// Storage offset of the FP32 state: the device holds u - off on the constant mode, so that
// point values are U = sum_i phi_i u_i + uref with uref = phi_0 off.  Rounding of the stored
// coefficients then scales with the deviation from uref instead of with |U|.  Zero in FP64.
struct Off4 {
    double v[4];
};

struct StageParams {
    const void* uk;
    const void* un;
    void* out;
    const double* geo;
    const int32_t* nbr;
    const uint32_t* code;
    const void* tab;        // tables in the state precision
    const void* tabm;       // lane-major tables of the MFMA kernel (FP32 Euler / Navier-Stokes), or null
    const void* tabt;       // face-trace and external-face tables of the MFMA Navier-Stokes kernel, or null
    const double* bc;       // boundary-face point rows (bc_row)
    const double* minve;    // quads: per-element inverse mass matrices [Np*Np][Kp] (FP64), else null
    const double* tab64;    // quads: the tables in FP64 (geometry nodes and weights, mass factors)
    int K, Kp;
    double dt, ca, cb;
    double gamma, mu, prandtl, sigma0;
    int mode;
    int ns_face_quirk;
    int ns_avg_ref;  // viscous face flux: 1 the reference's (half difference, inviscid sign), 0 consistent
    double kappa;    // heat coefficient of q = -kappa grad T in the energy flux u.tau + q: mu gamma /
                     // (Pr (gamma - 1)) as the reference, its negative for conduction
    double uref[4];  // point-value storage offset (see Off4)
    double uoff[4];  // the same as a constant-mode coefficient (uref / phi_0)
    // matrix-core tile kernels (null elsewhere): FP32 geometry of this mode, face connections,
    // external-face lists, and the largest count of interior external faces of a tile (sizes
    // the dynamic LDS of their traces)
    const float* geo32;
    const uint32_t* conn;
    const int32_t* ext;       // int2 entries
    const int32_t* ext_off;   // [tiles + 1]
    const int32_t* ext_nint;  // [tiles]
    const int32_t* halo;      // int2 [tiles][64]: the first 64 external faces with a neighbour of each
                              // tile (the ext entries, x = -1 past the tile's count)
    int halo_cap;
    // the FP32 tile kernels' scalars, rounded once on the host so that the kernels take them as
    // wave-uniform operands: gamma, mu, the viscous lambda -2 mu / 3 (FP32 division), kappa,
    // the penalty prefactor sigma0 mu (p+1)^2, dt, ca, cb, uref and uoff
    struct F32 {
        float gamma, mu, lam, kappa, sig0, dt, ca, cb, uref[4], uoff[4];
    } f;
};

// Kernel launchers (dgx_kernels.hip / dgx_kernels_ns.hip).  Each returns hipSuccess or the
// launch error.
hipError_t launch_stage(int precision, int physics, int p, int rule, const StageParams& sp,
                        hipStream_t s);
// pack/project subtract `off` from the constant mode, unpack adds it, eval adds `uref`.
hipError_t launch_project(int precision, int p, int rule, const double* Uq, void* u,
                          const void* tabL, int K, int Kp, const double* w_minv, Off4 off, hipStream_t s);
hipError_t launch_eval(int precision, int p, int rule, const void* u, const void* tab, int K,
                       int Kp, double* Uq, Off4 uref, hipStream_t s);
hipError_t launch_eval_n(int precision, int np, int nq, const void* u, const void* phiV64, int K, int Kp, double* Uq,
                         Off4 uref, hipStream_t s);
hipError_t launch_pack(int precision, int np, const double* host_layout_dev, void* u, int K,
                       int Kp, Off4 off, hipStream_t s);
hipError_t launch_unpack(int precision, int np, const void* u, double* host_layout_dev, int K,
                         int Kp, Off4 off, hipStream_t s);
// Per-block FP64 partial reductions: 0 conservation sums (4), 1 kinetic energy (1),
// 2 CFL/density/finite (max wave/h, rho min, rho max, bad count)
hipError_t launch_reduce(int kind, int precision, int p, int rule, const void* u, const void* tab,
                         const double* geo, const double* hmin, const double* qw, int K, int Kp,
                         double gamma, Off4 uref, Off4 off, double* partial, int nblocks, hipStream_t s);
// Relative-L2 error partial sums (dgx_aux.hip): 8 doubles per block of 256 elements.
hipError_t launch_geometry(const double* V, const int32_t* T, const int32_t* N, const int32_t* caller, int K, int Kp,
                           double* geo, double* hmin, uint32_t* code, int* err, hipStream_t s);
// per_point: the weights are the quad geometry rows w_q |J_q| instead of qw[q] |J|.
hipError_t launch_l2_error(const double* Uq, const double* Ex, int ex_prim, const double* geo, const double* qw, int nq,
                           int K, int Kp, double gamma, double* partial, int nblocks, int per_point, hipStream_t s);
// Quadrilateral elements (dgx_quad.hip).
hipError_t launch_stage_quad(int precision, int physics, int p, bool quirk, const StageParams& sp, hipStream_t s);
// Euler on quads, p <= 3, four lanes per element (dgx_quad_tensor.hip)
hipError_t launch_stage_quad_tensor(int precision, int p, const StageParams& sp, hipStream_t s);
hipError_t launch_project_quad(int precision, int np, int nq, const double* Uq, void* u, const double* phiV,
                               const double* geo, const double* minve, int K, int Kp, Off4 off, hipStream_t s);
hipError_t launch_minv_quad(int np, const double* minve, double* r, int K, int Kp, hipStream_t s);
hipError_t launch_reduce_quad(int kind, int precision, int np, int nq, const void* u, const double* phiV,
                              const double* geo, const double* hmin, int K, int Kp, double gamma, Off4 uref,
                              double* partial, int nblocks, hipStream_t s);
hipError_t launch_apply_minv(int precision, int np, const double* minv_ref, const double* geo,
                             double* r_dev, int K, int Kp, hipStream_t s);
// Parameters of one stage-adjoint launch (dgx_vjp.hip, dgx_vjp_mfma.hip).
struct VjpParams {
    const void* uk;      // stage input state
    const void* lam;     // adjoint of the stage output (or of L / R in RHS / RESIDUAL mode)
    const void* add;     // optional addend for the output (may be null)
    const void* acc_in;  // optional accumulator input (may be null)
    void* acc_out;       // optional accumulator output: acc_in + ca lam (may be null)
    void* out;           // lam_uk
    const double* geo;
    const int32_t* nbr;
    const uint32_t* code;
    const void* tab;
    const double* bc;
    int K, Kp;
    double dt, ca, cb, gamma;
    int mode;
    double uref[4];    // point-value storage offset of the state (zero in FP64)
    const void* tabm;  // lane tables of the matrix-core VJP kernel (FP32 Euler), or null
    // matrix-core tile kernel (null elsewhere): the forward tile kernel's FP32 geometry of this
    // mode, face connections and external-face lists (StageParams), and its scalars rounded to
    // FP32 on the host
    const float* geo32;
    const uint32_t* conn;
    const int32_t* ext;
    const int32_t* ext_off;
    const int32_t* ext_nint;
    const int32_t* halo;
    int halo_cap;
    struct F32 {
        float gamma, dt, ca, cb, uref[4];
    } f;
};

// Euler stage adjoint (dgx_vjp.hip): out = cb (lam + dt (dL/du)^T lam) + add,
// acc_out = acc_in + ca lam (mode RK); out = (dL/du)^T lam (mode RHS) or (dR/du)^T lam
// (mode RESIDUAL, unfolded tables).
hipError_t launch_vjp(int precision, int p, int rule, const VjpParams& vp, hipStream_t s);
// Batched persistent stage (dgx_batch.hip): items of (instance, first element, count, order);
// params[instance] holds that instance's stage parameters; tabs[p-1] the order-p tables.
struct BatchItem {
    int32_t instance, e0, count, order;
};
hipError_t launch_batch_stage(int precision, const StageParams* params, const BatchItem* items, int n_items,
                              const void* const tabs[4], int order_mask, int grid, hipStream_t s);
hipError_t launch_batch_stage_order(int precision, int p, const StageParams* params, const BatchItem* items,
                                    int n_items, const void* tab, int grid, hipStream_t s);
// Batched FP32 Euler items of order p = 3 or 4 on the matrix cores (dgx_stage_mfma.hip): block
// work items of whole 64-element tiles, `tab` the order's lane tables.
hipError_t launch_batch_stage_mfma(int p, const StageParams* params, const BatchItem* items, int n_items,
                                   int halo_cap, int grid, hipStream_t s);
// FP32 tile-kernel geometry (Geo32Row) from the FP64 geometry, for the RK/RHS tables (residual 0)
// or for M dU/dt (residual 1)
hipError_t launch_geometry32(const double* geo, int K, int Kp, int residual, float* out, hipStream_t s);
// gradient of the kinetic-energy functional with respect to the coefficients (device layout)
hipError_t launch_ke_grad(int precision, int p, int rule, const void* u, const double* tab64, const double* qw,
                          const double* geo, void* grad, int K, int Kp, Off4 uref, hipStream_t s);

}  // namespace dgx
