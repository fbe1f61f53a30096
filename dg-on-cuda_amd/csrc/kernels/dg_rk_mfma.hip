// Launcher of the matrix-core RK4 stage (kernels and tables in
// dg_stage_mfma.hpp): the FP32-contraction form and the FP64 matrix-core
// variant, both operators. The nontemporal-stream
// instantiations live in dg_rk_mfma_nt.hip.
#include "dg_stage_mfma.hpp"

namespace dgc {

hipError_t launch_stage_mfma(int order, int physics, int stage, const StageParams& p, hipStream_t s) {
  if (stage < 0 || stage > 4) return hipErrorInvalidValue;
  if (p.opF == nullptr && p.opD == nullptr) return hipErrorInvalidValue;
  const bool f32 = p.opF != nullptr;
  StageFn fn = p.streaming ? mfma_detail::pick_streaming(order, physics, f32, stage)
                           : mfma_detail::pick<false>(order, physics, f32, stage);
  if (fn == nullptr) return hipErrorInvalidValue;
  const int grid = (p.K + kTile - 1) / kTile;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(kBlock), 0, s, p);
  return hipGetLastError();
}

}  // namespace dgc
