// Stage-kernel dispatch over the (precision, physics) instantiation units.
#include "dgx_mfma.hpp"

namespace dgx {

hipError_t launch_stage_f32_euler(int p, int rule, bool quirk, const StageParams& sp, hipStream_t s);
hipError_t launch_stage_f64_euler(int p, int rule, bool quirk, const StageParams& sp, hipStream_t s);
hipError_t launch_stage_f32_ns(int p, int rule, bool quirk, const StageParams& sp, hipStream_t s);
hipError_t launch_stage_f64_ns(int p, int rule, bool quirk, const StageParams& sp, hipStream_t s);

hipError_t launch_stage_mfma(int p, bool bf16vol, const StageParams& sp, hipStream_t s);
hipError_t launch_stage_mfma_ns(int p, bool quirk, const StageParams& sp, hipStream_t s);

hipError_t launch_stage(int precision, int physics, int p, int rule, const StageParams& sp, hipStream_t s) {
    if (sp.K <= 0) return hipSuccess;
    // FP32 Euler with the degree-2p volume rule runs on the matrix cores when its lane tables
    // were built for the context
    if (sp.tabm && precision != DGX_FP64 && physics == DGX_EULER && rule == 0 && mfma_euler_order(p))
        return launch_stage_mfma(p, precision == DGX_FP32_BF16VOL, sp, s);
    const bool quirk = sp.ns_face_quirk != 0;
    // FP32 Navier-Stokes likewise (tile kernel with its own face-trace tables)
    if (sp.tabm && sp.tabt && sp.conn && precision == DGX_FP32 && physics == DGX_NAVIER_STOKES && rule == 0 &&
        mfma_order(p))
        return launch_stage_mfma_ns(p, quirk, sp, s);
    const bool f64 = precision == DGX_FP64;
    if (physics == DGX_EULER)
        return f64 ? launch_stage_f64_euler(p, rule, quirk, sp, s) : launch_stage_f32_euler(p, rule, quirk, sp, s);
    return f64 ? launch_stage_f64_ns(p, rule, quirk, sp, s) : launch_stage_f32_ns(p, rule, quirk, sp, s);
}

}  // namespace dgx
