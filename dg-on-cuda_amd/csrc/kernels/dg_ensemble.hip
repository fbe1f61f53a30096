// Launcher of the batched-ensemble stage kernels (dg_ensemble.hpp).
#include "dg_ensemble.hpp"

namespace dgc {

bool ensemble_supports(int order, int wt) { return ens_detail::ensemble_fn<false>(order, wt, 1) != nullptr; }

hipError_t launch_ensemble_stage(int order, int wt, bool streaming, int stage, const StageParams* params,
                                 const int2* items, int nitems, hipStream_t s) {
  ens_detail::EnsFn fn = streaming ? ens_detail::ensemble_fn_streaming(order, wt, stage)
                                   : ens_detail::ensemble_fn<false>(order, wt, stage);
  if (fn == nullptr || stage < 1 || stage > 4 || nitems < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(fn, dim3(nitems), dim3(group_threads(wt)), 0, s, params, items);
  return hipGetLastError();
}

}  // namespace dgc
