// Per-order kernels of the fused FP64 stage (device code in
// dg_stage_tile.hpp) and their selection tables. The two cache-hint variants
// are instantiated in separate translation units: dg_rk.hip (temporal) and
// dg_rk_nt.hip (nontemporal accumulator streams).
#pragma once
#include "dg_stage_tile.hpp"

namespace dgc {

template <int N, int PHYS, int STAGE, int WT, bool NT>
__global__ __launch_bounds__(group_threads(WT)) void dg_stage_kernel(const StageParams p) {
  __shared__ double lds[stage_lds_doubles(N, PHYS, WT)];
  stage_tile<N, PHYS, STAGE, WT, NT>(p, xcd_group(blockIdx.x, gridDim.x), lds);
}

namespace rk_detail {

template <int N, int PHYS, int WT, bool NT>
constexpr StageFn stage_fn(int stage) {
  return stage == 0 ? dg_stage_kernel<N, PHYS, 0, WT, NT>
       : stage == 1 ? dg_stage_kernel<N, PHYS, 1, WT, NT>
       : stage == 2 ? dg_stage_kernel<N, PHYS, 2, WT, NT>
       : stage == 3 ? dg_stage_kernel<N, PHYS, 3, WT, NT>
                    : dg_stage_kernel<N, PHYS, 4, WT, NT>;
}

// A variant is built only if its LDS (all tiles of the workgroup) stays
// within 80 KB, so that at least two workgroups share a CU; larger requests
// fall back to one tile per workgroup (WT = 4). WT = 8, 16 and the split
// tiling hold one tile.
template <int N, int PHYS, int WT>
constexpr bool fits() {
  return WT >= 4 || stage_lds_doubles(N, PHYS, WT) * 8 <= 80 * 1024;
}

template <int N, int PHYS, bool NT>
StageFn pick_wt(int wt, int stage) {
  if constexpr (fits<N, PHYS, 1>()) {
    if (wt == 1) return stage_fn<N, PHYS, 1, NT>(stage);
  }
  if constexpr (fits<N, PHYS, 2>()) {
    if (wt == 2) return stage_fn<N, PHYS, 2, NT>(stage);
  }
  if (wt == 8) return stage_fn<N, PHYS, 8, NT>(stage);
  if (wt == 16) return stage_fn<N, PHYS, 16, NT>(stage);
  if (wt == kSplitTiling) return stage_fn<N, PHYS, kSplitTiling, NT>(stage);
  return stage_fn<N, PHYS, 4, NT>(stage);
}

template <int PHYS, bool NT>
StageFn pick_order(int order, int wt, int stage) {
  switch (order) {
    case 1: return pick_wt<1, PHYS, NT>(wt, stage);
    case 2: return pick_wt<2, PHYS, NT>(wt, stage);
    case 3: return pick_wt<3, PHYS, NT>(wt, stage);
    case 4: return pick_wt<4, PHYS, NT>(wt, stage);
    case 5: return pick_wt<5, PHYS, NT>(wt, stage);
    case 6: return pick_wt<6, PHYS, NT>(wt, stage);
    case 7: return pick_wt<7, PHYS, NT>(wt, stage);
    default: return nullptr;
  }
}

template <bool NT>
StageFn pick(int order, int physics, int wt, int stage) {
  return physics == kMaxwellTM ? pick_order<kMaxwellTM, NT>(order, wt, stage)
                               : pick_order<kAdvection, NT>(order, wt, stage);
}

// The nontemporal table (dg_rk_nt.hip).
StageFn pick_streaming(int order, int physics, int wt, int stage);

}  // namespace rk_detail
}  // namespace dgc
