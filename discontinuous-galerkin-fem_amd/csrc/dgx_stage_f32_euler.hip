// Stage-kernel instantiations: float state, Euler.
#include "dgx_stage_impl.hpp"

namespace dgx {
hipError_t launch_stage_f32_euler(int p, int rule, bool quirk, const StageParams& sp, hipStream_t s) {
    (void)quirk;
    return launch_family<float, false, false>(p, rule, sp, s);
}
}  // namespace dgx
