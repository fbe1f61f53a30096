// Per-order launchers of the matrix-core RK4 stage (device code in
// dg_stage_mfma.hpp): the FP32-contraction form of the advection operator and
// the FP64 matrix-core variant of both operators.
#include "dg_stage_mfma.hpp"

namespace dgc {

namespace {

// Workgroup renumbering so that each of the 8 XCDs sweeps a contiguous range
// of tiles (same mapping as the VALU stage kernels).
__device__ __forceinline__ int xcd_tile(int b, int G) {
  constexpr int X = 8;
  const int x = b % X, i = b / X, per = G / X, rem = G % X;
  return x * per + (x < rem ? x : rem) + i;
}

template <typename CT, int N, int PHYS, int STAGE>
__global__ __launch_bounds__(kBlock) void dg_stage_mfma_kernel(const StageParams p) {
  __shared__ __align__(16) unsigned char lds[mfma_lds_bytes<CT, N, PHYS>()];
  stage_tile_mfma<CT, N, PHYS, STAGE>(p, xcd_tile(blockIdx.x, gridDim.x), lds);
}

using StageFn = void (*)(const StageParams);

template <typename CT, int N, int PHYS>
StageFn mfma_fn(int stage) {
  return stage == 0 ? dg_stage_mfma_kernel<CT, N, PHYS, 0>
       : stage == 1 ? dg_stage_mfma_kernel<CT, N, PHYS, 1>
       : stage == 2 ? dg_stage_mfma_kernel<CT, N, PHYS, 2>
       : stage == 3 ? dg_stage_mfma_kernel<CT, N, PHYS, 3>
                    : dg_stage_mfma_kernel<CT, N, PHYS, 4>;
}

template <typename CT, int PHYS>
StageFn pick_mfma(int order, int stage) {
  switch (order) {
    case 1: return mfma_fn<CT, 1, PHYS>(stage);
    case 2: return mfma_fn<CT, 2, PHYS>(stage);
    case 3: return mfma_fn<CT, 3, PHYS>(stage);
    case 4: return mfma_fn<CT, 4, PHYS>(stage);
    case 5: return mfma_fn<CT, 5, PHYS>(stage);
    case 6: return mfma_fn<CT, 6, PHYS>(stage);
    case 7: return mfma_fn<CT, 7, PHYS>(stage);
    default: return nullptr;
  }
}

}  // namespace

hipError_t launch_stage_mfma(int order, int physics, int stage, const StageParams& p, hipStream_t s) {
  if (stage < 0 || stage > 4) return hipErrorInvalidValue;
  StageFn fn = nullptr;
  if (p.opF != nullptr) {
    // the FP32-contraction form is built for the advection operator
    if (physics == kAdvection) fn = pick_mfma<float, kAdvection>(order, stage);
  } else if (p.opD != nullptr) {
    fn = physics == kMaxwellTM ? pick_mfma<double, kMaxwellTM>(order, stage)
                               : pick_mfma<double, kAdvection>(order, stage);
  }
  if (fn == nullptr) return hipErrorInvalidValue;
  const int grid = (p.K + kTile - 1) / kTile;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(kBlock), 0, s, p);
  return hipGetLastError();
}

}  // namespace dgc
