// Instantiations of the matrix-core stage kernels with nontemporal
// accumulator streams (selected by launch_stage_mfma in dg_rk_mfma.hip).
#include "dg_stage_mfma.hpp"

namespace dgc::mfma_detail {

StageFn pick_streaming(int order, int physics, bool f32, int stage) {
  return pick<true>(order, physics, f32, stage);
}

}  // namespace dgc::mfma_detail
