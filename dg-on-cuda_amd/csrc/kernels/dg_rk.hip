// Launcher of the fused FP64 stage kernels (dg_rk_kernels.hpp; device code
// in dg_stage_tile.hpp, shared with the ensemble kernel).
#include "dg_rk_kernels.hpp"

namespace dgc {

namespace {

using rk_detail::fits;

template <int N, int PHYS>
int effective_wt(int wt) {
  if (wt == 1 && fits<N, PHYS, 1>()) return 1;
  if (wt == 2 && fits<N, PHYS, 2>()) return 2;
  if (wt == 8 || wt == 16 || wt == kSplitTiling) return wt;
  return 4;
}

}  // namespace

// The waves-per-tile value (or kMfmaTiling) that will actually run for a request.
int supported_waves_per_tile(int order, int physics, int wt) {
  if (wt == kMfmaTiling) return order >= 1 && order <= kMaxOrder ? kMfmaTiling : 4;
#define DGC_WT_CASE(n) \
  case n: return physics == kMaxwellTM ? effective_wt<n, kMaxwellTM>(wt) : effective_wt<n, kAdvection>(wt);
  switch (order) {
    DGC_WT_CASE(1)
    DGC_WT_CASE(2)
    DGC_WT_CASE(3)
    DGC_WT_CASE(4)
    DGC_WT_CASE(5)
    DGC_WT_CASE(6)
    DGC_WT_CASE(7)
    default: return 4;
  }
#undef DGC_WT_CASE
}

int default_waves_per_tile(int order, int physics, int K) {
  // Meshes of a few tiles leave most of the GPU idle and each stage is one
  // dependent chain of loads and products per tile: more waves per tile, or
  // a tile's rows split over four workgroups, shorten it. Fastest whole RK4
  // step per operator and order, measured at K = 512, 2048 and 4050 (the
  // same choice at all three; Maxwell K = 512, N = 4: 47.9 us per step with
  // 4 waves per tile, 26.1 with 16, 24.4 split).
  if ((K + kTile - 1) / kTile <= 64) {
    static constexpr int kMaxwell[kMaxOrder + 1] = {0, 8, 16, 16, kSplitTiling, 16, 16, 16};
    static constexpr int kAdv[kMaxOrder + 1] = {0, 8, 8, kSplitTiling, kSplitTiling, 8, kMfmaTiling, 16};
    if (order >= 1 && order <= kMaxOrder) return physics == kMaxwellTM ? kMaxwell[order] : kAdv[order];
    return 16;
  }
  const int want = physics == kAdvection ? (order <= 1 ? 1 : (order <= 3 ? 2 : 4)) : (order <= 1 ? 2 : 4);
  return supported_waves_per_tile(order, physics, want);
}

hipError_t launch_stage(int order, int physics, int stage, int wt, const StageParams& p, hipStream_t s) {
  if (p.opF != nullptr) return launch_stage_mfma(order, physics, stage, p, s);
  if (wt == kMfmaTiling) {
    if (p.opD == nullptr) return hipErrorInvalidValue;
    return launch_stage_mfma(order, physics, stage, p, s);
  }
  if (stage < 0 || stage > 4 || (wt != 1 && wt != 2 && wt != 4 && wt != 8 && wt != 16 && wt != kSplitTiling))
    return hipErrorInvalidValue;
  wt = supported_waves_per_tile(order, physics, wt);
  StageFn fn = p.streaming ? rk_detail::pick_streaming(order, physics, wt, stage)
                           : rk_detail::pick<false>(order, physics, wt, stage);
  if (fn == nullptr) return hipErrorInvalidValue;
  const int tiles = (p.K + kTile - 1) / kTile;
  const int tpw = tiles_per_group(wt);
  const int grid = (tiles + tpw - 1) / tpw * tile_splits(wt);
  hipLaunchKernelGGL(fn, dim3(grid), dim3(group_threads(wt)), 0, s, p);
  return hipGetLastError();
}

}  // namespace dgc
