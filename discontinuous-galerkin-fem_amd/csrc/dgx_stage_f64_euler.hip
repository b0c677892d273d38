// Stage-kernel instantiations: double state, Euler.
#include "dgx_stage_impl.hpp"

namespace dgx {
hipError_t launch_stage_f64_euler(int p, int rule, bool quirk, const StageParams& sp, hipStream_t s) {
    (void)quirk;
    return launch_family<double, false, false>(p, rule, sp, s);
}
}  // namespace dgx
