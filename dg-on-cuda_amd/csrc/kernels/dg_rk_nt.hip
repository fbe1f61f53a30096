// Instantiations of the per-order stage kernels with nontemporal accumulator
// streams (selected by launch_stage in dg_rk.hip).
#include "dg_rk_kernels.hpp"

namespace dgc::rk_detail {

StageFn pick_streaming(int order, int physics, int wt, int stage) { return pick<true>(order, physics, wt, stage); }

}  // namespace dgc::rk_detail
