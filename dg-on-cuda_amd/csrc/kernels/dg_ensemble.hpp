// Batched ensemble: many independent Maxwell TM instances (own mesh, size
// and order each) advance together, one RK4 stage at a time.
//
// One launch per order and stage covers every instance of that order: its
// work list holds one item per (instance, group of 4/WT tiles of 64 cells;
// one tile at WT = 8),
// the items of an instance contiguous, and each XCD takes a contiguous range
// of the list, so the neighbour traces a tile reads were usually loaded into
// the same L2 by the tile before it. Per item the workgroup runs the device
// code of the per-order kernels (stage_tile, dg_stage_tile.hpp) with that
// instance's stage parameters. Kernels of different orders share nothing, so
// the step graph runs the orders as independent branches.
// No reference counterpart: the reference runs one mesh per process.
//
// The temporal variants are instantiated in dg_ensemble.hip, the variants
// with nontemporal accumulator streams in dg_ensemble_nt.hip.
#pragma once
#include <cstdint>

#include "dg_stage_mfma.hpp"
#include "dg_stage_tile.hpp"

namespace dgc {

namespace ens_detail {

// The per-(instance, stage) parameter records are read through the constant
// address space: like kernel arguments they are then invariant scalar loads
// that the compiler issues next to their uses instead of holding them in
// registers for the whole item.
using ConstParams = const StageParams __attribute__((address_space(4)));

// One workgroup per item, items mapped so that each XCD takes a contiguous
// range. (A persistent variant looping over items let the compiler hoist
// per-item address arithmetic out of the loop: 191 -> 256 VGPRs plus AGPR
// spills at N = 3, half the occupancy.)
template <int N, int WT, int STAGE, bool NT>
__global__ __launch_bounds__(group_threads(WT)) void ensemble_stage_kernel(const StageParams* __restrict__ params,
                                                                const int2* __restrict__ items) {
  __shared__ double lds[stage_lds_doubles(N, kMaxwellTM, WT)];
  ConstParams* base = reinterpret_cast<ConstParams*>(reinterpret_cast<uintptr_t>(params));
  const int2 item = items[xcd_group(blockIdx.x, gridDim.x)];  // (instance, tile group)
  const StageParams& p = *(const StageParams*)(base + (uniform(item.x) * 4 + STAGE - 1));
  stage_tile<N, kMaxwellTM, STAGE, WT, NT>(p, uniform(item.y), lds);
}

// The FP64 matrix-core variant (tiling 0): one 64-cell tile per item.
template <int N, int STAGE, bool NT>
__global__ __launch_bounds__(kBlock) void ensemble_mfma_kernel(const StageParams* __restrict__ params,
                                                               const int2* __restrict__ items) {
  __shared__ __align__(16) unsigned char lds[mfma_lds_bytes<double, N, kMaxwellTM>()];
  ConstParams* base = reinterpret_cast<ConstParams*>(reinterpret_cast<uintptr_t>(params));
  const int2 item = items[xcd_group(blockIdx.x, gridDim.x)];
  const StageParams& p = *(const StageParams*)(base + (uniform(item.x) * 4 + STAGE - 1));
  stage_tile_mfma<double, N, kMaxwellTM, STAGE, NT>(p, uniform(item.y), lds);
}

using EnsFn = void (*)(const StageParams*, const int2*);

template <int N, int WT, bool NT>
EnsFn stage_of(int stage) {
  if constexpr (WT == kMfmaTiling) {
    return stage == 1 ? ensemble_mfma_kernel<N, 1, NT>
         : stage == 2 ? ensemble_mfma_kernel<N, 2, NT>
         : stage == 3 ? ensemble_mfma_kernel<N, 3, NT>
                      : ensemble_mfma_kernel<N, 4, NT>;
  } else {
    return stage == 1 ? ensemble_stage_kernel<N, WT, 1, NT>
         : stage == 2 ? ensemble_stage_kernel<N, WT, 2, NT>
         : stage == 3 ? ensemble_stage_kernel<N, WT, 3, NT>
                      : ensemble_stage_kernel<N, WT, 4, NT>;
  }
}

template <bool NT>
EnsFn ensemble_fn(int order, int wt, int stage) {
  if (wt == 8) {  // one tile per 512-thread workgroup
    switch (order) {
      case 1: return stage_of<1, 8, NT>(stage);
      case 2: return stage_of<2, 8, NT>(stage);
      case 3: return stage_of<3, 8, NT>(stage);
      case 4: return stage_of<4, 8, NT>(stage);
      case 5: return stage_of<5, 8, NT>(stage);
      case 6: return stage_of<6, 8, NT>(stage);
      case 7: return stage_of<7, 8, NT>(stage);
      default: return nullptr;
    }
  }
  if (wt != kMfmaTiling && wt != 1 && wt != 2 && wt != 4) return nullptr;
  switch (order * 8 + wt) {
    case 1 * 8 + 2: return stage_of<1, 2, NT>(stage);
    case 1 * 8 + 4: return stage_of<1, 4, NT>(stage);
    case 2 * 8 + 2: return stage_of<2, 2, NT>(stage);
    case 2 * 8 + 4: return stage_of<2, 4, NT>(stage);
    case 3 * 8 + 2: return stage_of<3, 2, NT>(stage);
    case 3 * 8 + 4: return stage_of<3, 4, NT>(stage);
    case 4 * 8 + 4: return stage_of<4, 4, NT>(stage);
    case 5 * 8 + 4: return stage_of<5, 4, NT>(stage);
    case 6 * 8 + 4: return stage_of<6, 4, NT>(stage);
    case 7 * 8 + 4: return stage_of<7, 4, NT>(stage);
    case 1 * 8 + kMfmaTiling: return stage_of<1, kMfmaTiling, NT>(stage);
    case 2 * 8 + kMfmaTiling: return stage_of<2, kMfmaTiling, NT>(stage);
    case 3 * 8 + kMfmaTiling: return stage_of<3, kMfmaTiling, NT>(stage);
    case 4 * 8 + kMfmaTiling: return stage_of<4, kMfmaTiling, NT>(stage);
    case 5 * 8 + kMfmaTiling: return stage_of<5, kMfmaTiling, NT>(stage);
    case 6 * 8 + kMfmaTiling: return stage_of<6, kMfmaTiling, NT>(stage);
    case 7 * 8 + kMfmaTiling: return stage_of<7, kMfmaTiling, NT>(stage);
    default: return nullptr;
  }
}

// The nontemporal table (dg_ensemble_nt.hip).
EnsFn ensemble_fn_streaming(int order, int wt, int stage);

}  // namespace ens_detail
}  // namespace dgc
