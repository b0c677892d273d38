// Stage-kernel instantiations: float state, Navier-Stokes (both face-gradient reads).
#include "dgx_stage_impl.hpp"

namespace dgx {
hipError_t launch_stage_f32_ns(int p, int rule, bool quirk, const StageParams& sp, hipStream_t s) {
    (void)quirk;
    return quirk ? launch_family<float, true, true>(p, rule, sp, s) : launch_family<float, true, false>(p, rule, sp, s);
}
}  // namespace dgx
