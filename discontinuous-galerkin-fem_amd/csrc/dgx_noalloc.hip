// dgx_conservation_sums_device: the conservation sums of a context's state written to device
// memory, with no allocation, no host synchronisation and no memory set, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops).  The per-block FP64
// partial sums that dgx_conservation_sums keeps in the context's buffer live here in one
// file-scope device array per GPU.
#include <mutex>

#include "dgx_ctx.hpp"

namespace {

// 32768 blocks x 4 doubles = 1 MiB: contexts of up to 32768 * 256 = 8388608 elements.
constexpr int kScratchBlocks = 32768;
constexpr int kRedElems = 256;  // elements per partial (the reduction kernel's block)
static __device__ __align__(256) double scratch[kScratchBlocks * 4];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on device `dev`: looked up once per device with that device
// current (each GPU holds its own copy of the array).  In host code the array's name is not a
// device address; only this pointer is passed to kernels.
double* scratch_on(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw Fail(DGX_E_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        hip_check(hipGetDevice(&cur), "hipGetDevice");
        if (cur != dev) hip_check(hipSetDevice(dev), "hipSetDevice");
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        hip_check(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

// out[v] = sum over the nblocks partials of variable v in block order: the same sequence of
// FP64 additions as dgx_conservation_sums performs on the host, so the two agree bit for bit.
__global__ __launch_bounds__(64) void sum_partials_kernel(const double* __restrict__ partial, int nblocks,
                                                          double* __restrict__ out) {
    const int v = threadIdx.x;
    if (v >= 4) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += partial[b * 4 + v];
    out[v] = s;
}

}  // namespace

extern "C" dgx_status dgx_conservation_sums_device(dgx_ctx* c, double* out_dev) {
    return guarded([&] {
        if (!c || !out_dev) throw std::invalid_argument("null argument");
        const int nb = c->red_blocks();
        if (nb > kScratchBlocks)
            throw Fail(DGX_E_UNSUPPORTED, "dgx_conservation_sums_device: the mesh has " + std::to_string(c->K) +
                                              " elements, more than the " + std::to_string(kScratchBlocks * kRedElems) +
                                              " its fixed device scratch covers; use dgx_conservation_sums");
        double* part = scratch_on(c->device);
        // every partial read below is written by this launch: nothing needs clearing
        if (c->nf == 4)
            hip_check(launch_reduce_quad(0, c->precision, c->np, c->nq, c->u.p, c->tab64.as<double>(),
                                         c->geo.as<double>(), c->hmin.as<double>(), c->K, c->Kp, c->gamma, c->uref,
                                         part, nb, c->stream),
                      "reduce kernel launch");
        else
            hip_check(launch_reduce(0, c->precision, c->p, c->rule, c->u.p, c->tab64.p, c->geo.as<double>(),
                                    c->hmin.as<double>(), c->qw.as<double>(), c->K, c->Kp, c->gamma, c->uref, c->off,
                                    part, nb, c->stream),
                      "reduce kernel launch");
        hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(64), 0, c->stream, part, nb, out_dev);
        hip_check(hipGetLastError(), "sum kernel launch");
    });
}
