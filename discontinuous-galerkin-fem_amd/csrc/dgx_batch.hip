// Batched, persistent SSP-RK3 stage over many independent Euler instances of mixed mesh size
// and polynomial order (FP32 or FP64).  One launch per RK stage covers every instance: the
// work list holds items (instance, first element, element count, order); each wave walks the
// list with a fixed stride (no atomics, so the launches can be captured in a graph and
// replayed), and the element work is the same fused stage as the single-instance kernel.
// Per-instance launch parameters (state pointers, dt, geometry) live in device memory.
#include "dgx_stage_impl.hpp"

namespace dgx {
namespace {

constexpr int kWaves = kBlock / 64;

template <typename Real, int P>
__device__ __forceinline__ void run_item(const StageParams& sp, int e0, int count, int lane, const Real* lds) {
    using T = StageTab<P, 0, false>;
    const Real* s_phiF = lds;
    const Real* s_phiFT = lds + T::NF;
    for (int e = e0 + lane; e < e0 + count; e += 64)
        stage_element<P, 0, Real, false, false>(sp, e, sp.dt, s_phiF, s_phiFT, nullptr);
}

// LDS offsets of the neighbour tables of orders 1..4 (two copies of 3 NFQ NP each).
__host__ __device__ constexpr int lds_off(int p) {
    int o = 0;
    for (int q = 1; q < p; ++q) o += 2 * 3 * (q + 1) * np_of(q);
    return o;
}
constexpr int kLdsTotal = lds_off(5);

template <typename Real>
__global__ __launch_bounds__(kBlock) void batch_stage_kernel(const StageParams* __restrict__ params,
                                                             const BatchItem* __restrict__ items, int n_items,
                                                             const Real* tab1, const Real* tab2, const Real* tab3,
                                                             const Real* tab4, int order_mask) {
    __shared__ Real lds[kLdsTotal];
    if (order_mask & 1) stage_tables_to_lds<1, 0, Real, false>(tab1, lds + lds_off(1), lds + lds_off(1) + 3 * 2 * 3, nullptr, threadIdx.x, kBlock);
    if (order_mask & 2) stage_tables_to_lds<2, 0, Real, false>(tab2, lds + lds_off(2), lds + lds_off(2) + 3 * 3 * 6, nullptr, threadIdx.x, kBlock);
    if (order_mask & 4) stage_tables_to_lds<3, 0, Real, false>(tab3, lds + lds_off(3), lds + lds_off(3) + 3 * 4 * 10, nullptr, threadIdx.x, kBlock);
    if (order_mask & 8) stage_tables_to_lds<4, 0, Real, false>(tab4, lds + lds_off(4), lds + lds_off(4) + 3 * 5 * 15, nullptr, threadIdx.x, kBlock);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    // the wave index is wave-uniform: readfirstlane keeps the item and parameter loads scalar
    const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * kWaves + (threadIdx.x >> 6));
    const int n_waves = gridDim.x * kWaves;
    for (int it = wave; it < n_items; it += n_waves) {
        const BatchItem item = items[it];
        const StageParams sp = params[item.instance];
        switch (item.order) {
        case 1: run_item<Real, 1>(sp, item.e0, item.count, lane, lds + lds_off(1)); break;
        case 2: run_item<Real, 2>(sp, item.e0, item.count, lane, lds + lds_off(2)); break;
        case 3: run_item<Real, 3>(sp, item.e0, item.count, lane, lds + lds_off(3)); break;
        case 4: run_item<Real, 4>(sp, item.e0, item.count, lane, lds + lds_off(4)); break;
        default: break;
        }
    }
}

// One order only: the same walk without the order switch, so the kernel keeps that order's
// register budget (the "per-order" layout).
template <typename Real, int P>
__global__ __launch_bounds__(kBlock) void batch_stage_kernel_p(const StageParams* __restrict__ params,
                                                               const BatchItem* __restrict__ items, int n_items,
                                                               const Real* tab) {
    using T = StageTab<P, 0, false>;
    __shared__ Real s_phiF[T::NF];
    __shared__ Real s_phiFT[T::NF];
    stage_tables_to_lds<P, 0, Real, false>(tab, s_phiF, s_phiFT, nullptr, threadIdx.x, kBlock);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * kWaves + (threadIdx.x >> 6));
    const int n_waves = gridDim.x * kWaves;
    for (int it = wave; it < n_items; it += n_waves) {
        const BatchItem item = items[it];
        const StageParams sp = params[item.instance];
        for (int e = item.e0 + lane; e < item.e0 + item.count; e += 64)
            stage_element<P, 0, Real, false, false>(sp, e, sp.dt, s_phiF, s_phiFT, nullptr);
    }
}

}  // namespace

hipError_t launch_batch_stage(int precision, const StageParams* params, const BatchItem* items, int n_items,
                              const void* const tabs[4], int order_mask, int grid, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    if (precision == DGX_FP64)
        hipLaunchKernelGGL(batch_stage_kernel<double>, dim3(grid), dim3(kBlock), 0, s, params, items, n_items,
                           (const double*)tabs[0], (const double*)tabs[1], (const double*)tabs[2], (const double*)tabs[3],
                           order_mask);
    else
        hipLaunchKernelGGL(batch_stage_kernel<float>, dim3(grid), dim3(kBlock), 0, s, params, items, n_items,
                           (const float*)tabs[0], (const float*)tabs[1], (const float*)tabs[2], (const float*)tabs[3],
                           order_mask);
    return hipGetLastError();
}

hipError_t launch_batch_stage_order(int precision, int p, const StageParams* params, const BatchItem* items,
                                    int n_items, const void* tab, int grid, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
#define DGX_BP(PP)                                                                                          \
    if (precision == DGX_FP64)                                                                              \
        hipLaunchKernelGGL((batch_stage_kernel_p<double, PP>), dim3(grid), dim3(kBlock), 0, s, params, items, \
                           n_items, (const double*)tab);                                                    \
    else                                                                                                    \
        hipLaunchKernelGGL((batch_stage_kernel_p<float, PP>), dim3(grid), dim3(kBlock), 0, s, params, items,  \
                           n_items, (const float*)tab);
    switch (p) {
    case 1: DGX_BP(1) break;
    case 2: DGX_BP(2) break;
    case 3: DGX_BP(3) break;
    case 4: DGX_BP(4) break;
    default: return hipErrorInvalidValue;
    }
#undef DGX_BP
    return hipGetLastError();
}

}  // namespace dgx
