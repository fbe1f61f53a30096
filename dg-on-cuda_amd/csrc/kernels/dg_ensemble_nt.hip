// Instantiations of the ensemble stage kernels with nontemporal accumulator
// streams (selected by launch_ensemble_stage in dg_ensemble.hip).
#include "dg_ensemble.hpp"

namespace dgc::ens_detail {

EnsFn ensemble_fn_streaming(int order, int wt, int stage) { return ensemble_fn<true>(order, wt, stage); }

}  // namespace dgc::ens_detail
