// Per-order launchers of the fused FP64 stage (the device code is in
// dg_stage_tile.hpp, shared with the ensemble kernel).
#include "dg_stage_tile.hpp"

namespace dgc {

template <int N, int PHYS, int STAGE, int WT>
__global__ __launch_bounds__(kBlock) void dg_stage_kernel(const StageParams p) {
  __shared__ double lds[stage_lds_doubles(N, PHYS, WT)];
  stage_tile<N, PHYS, STAGE, WT>(p, xcd_group(blockIdx.x, gridDim.x), lds);
}

namespace {

using StageFn = void (*)(const StageParams);

template <int N, int PHYS, int WT>
constexpr StageFn stage_fn(int stage) {
  return stage == 0 ? dg_stage_kernel<N, PHYS, 0, WT>
       : stage == 1 ? dg_stage_kernel<N, PHYS, 1, WT>
       : stage == 2 ? dg_stage_kernel<N, PHYS, 2, WT>
       : stage == 3 ? dg_stage_kernel<N, PHYS, 3, WT>
                    : dg_stage_kernel<N, PHYS, 4, WT>;
}

// A variant is built only if its LDS (all tiles of the workgroup) stays
// within 80 KB, so that at least two workgroups share a CU; larger requests
// fall back to one tile per workgroup (WT = 4).
template <int N, int PHYS, int WT>
constexpr bool fits() {
  constexpr int Np = Tri<N>::Np, NF = 3 * Tri<N>::Nfp, F = PHYS == kMaxwellTM ? 3 : 1;
  return WT == 4 || (kWaves / WT) * F * (Np + NF) * kTile * 8 <= 80 * 1024;
}

template <int N, int PHYS>
StageFn pick_wt(int wt, int stage) {
  if constexpr (fits<N, PHYS, 1>()) {
    if (wt == 1) return stage_fn<N, PHYS, 1>(stage);
  }
  if constexpr (fits<N, PHYS, 2>()) {
    if (wt == 2) return stage_fn<N, PHYS, 2>(stage);
  }
  return stage_fn<N, PHYS, 4>(stage);
}

template <int N, int PHYS>
int effective_wt(int wt) {
  if (wt == 1 && fits<N, PHYS, 1>()) return 1;
  if (wt == 2 && fits<N, PHYS, 2>()) return 2;
  return 4;
}

template <int PHYS>
StageFn pick(int order, int wt, int stage) {
  switch (order) {
    case 1: return pick_wt<1, PHYS>(wt, stage);
    case 2: return pick_wt<2, PHYS>(wt, stage);
    case 3: return pick_wt<3, PHYS>(wt, stage);
    case 4: return pick_wt<4, PHYS>(wt, stage);
    case 5: return pick_wt<5, PHYS>(wt, stage);
    case 6: return pick_wt<6, PHYS>(wt, stage);
    case 7: return pick_wt<7, PHYS>(wt, stage);
    default: return nullptr;
  }
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

int default_waves_per_tile(int order, int physics) {
  const int want = physics == kAdvection ? (order <= 1 ? 1 : (order <= 3 ? 2 : 4)) : (order <= 1 ? 2 : 4);
  return supported_waves_per_tile(order, physics, want);
}

hipError_t launch_stage(int order, int physics, int stage, int wt, const StageParams& p, hipStream_t s) {
  if (p.opF != nullptr) return physics == kAdvection ? launch_stage_mfma(order, physics, stage, p, s) : hipErrorInvalidValue;
  if (wt == kMfmaTiling) {
    if (p.opD == nullptr) return hipErrorInvalidValue;
    return launch_stage_mfma(order, physics, stage, p, s);
  }
  if (stage < 0 || stage > 4 || (wt != 1 && wt != 2 && wt != 4)) return hipErrorInvalidValue;
  wt = supported_waves_per_tile(order, physics, wt);
  StageFn fn = physics == kMaxwellTM ? pick<kMaxwellTM>(order, wt, stage) : pick<kAdvection>(order, wt, stage);
  if (fn == nullptr) return hipErrorInvalidValue;
  const int tiles = (p.K + kTile - 1) / kTile;
  const int tpw = kWaves / wt;
  const int grid = (tiles + tpw - 1) / tpw;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(kBlock), 0, s, p);
  return hipGetLastError();
}

}  // namespace dgc
