// FP32 cell-relative form (placeholder until the kernels land).
#include "context.hpp"

namespace pdp3 {
void launch_fused_f32(Ctx &, Species &) {}
void launch_charge_weighting_f32(Ctx &, Species &) {}
void launch_bunch_inject_f32(Ctx &, Species &, unsigned, unsigned, double) {}
void launch_f32_from_f64(Ctx &, Species &) {}
void launch_f64_from_f32(Ctx &, Species &) {}
}  // namespace pdp3
