// Stage-kernel instantiations: double state, Navier-Stokes (both face-gradient reads).
#include "dgx_stage_impl.hpp"

namespace dgx {
hipError_t launch_stage_f64_ns(int p, int rule, bool quirk, const StageParams& sp, hipStream_t s) {
    (void)quirk;
    return quirk ? launch_family<double, true, true>(p, rule, sp, s) : launch_family<double, true, false>(p, rule, sp, s);
}
}  // namespace dgx
