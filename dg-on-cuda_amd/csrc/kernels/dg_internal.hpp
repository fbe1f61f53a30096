// Library-internal handle types and launch helpers shared by the HIP
// translation units behind the C ABI (include/dgc_hip.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <vector>

#include "dg_common.hpp"

struct dgc_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
};

// Device copy of one discretization (mirror of d_simple_discretization_2d).
struct dgc_disc_s {
  dgc_ctx_s* ctx = nullptr;
  int K = 0;
  int order = 0;
  int Np = 0, Nfp = 0;
  int num_boundary_nodes = 0;
  double* geo = nullptr;     // [kGeoComps][K]
  int* nbr = nullptr;        // [3][K] packed connectivity
  double* J = nullptr;       // [K]
  double* face_J = nullptr;  // [3][K]
  double* bx = nullptr;      // [num_boundary_nodes]
  double* by = nullptr;
  // host copies kept for set-up work that needs them (e.g. penalty factors)
  std::vector<double> h_J, h_face_J;
  std::vector<int> h_nbr;
};

namespace dgc {

// Two HIP events (start, stop) destroyed on every exit path of the timed
// entry points.
struct EventPair {
  hipEvent_t start = nullptr, stop = nullptr;
  hipError_t create() {
    hipError_t e = hipEventCreate(&start);
    return e == hipSuccess ? hipEventCreate(&stop) : e;
  }
  ~EventPair() {
    if (start) (void)hipEventDestroy(start);
    if (stop) (void)hipEventDestroy(stop);
  }
};

enum Physics { kMaxwellTM = 0, kAdvection = 1, kPhysicsCount = 2 };
enum AdvectionBC { kBcSinSin = 0, kBcZero = 1 };

// One fused RK4 stage (or a plain RHS evaluation when stage == 0).
struct StageParams {
  int K;
  int bc;
  const double* __restrict__ DrT;   // [Np][Npad] transposed, rows padded
  const double* __restrict__ DsT;   // [Np][Npad]
  const double* __restrict__ LT;    // [3Nfp][Npad]
  const double* __restrict__ geo;
  const int* __restrict__ nbr;
  const double* __restrict__ bx;
  const double* __restrict__ by;
  const double* __restrict__ tclock;  // device time of the step start (nullptr: use t)
  double ax, ay;
  double t;       // stage time when tclock is null
  double tfrac;   // stage time = tclock[0] + tfrac * dt_clock
  double dt_clock;
  double c1;      // stage coefficient (dt/2, dt/2, dt, dt/6)
  const double* in[3];
  double* out[3];
  double* u[3];
  double* acc[3];
  double* k3[3];
  // Matrix-core operand blocks of Dr, Ds and LIFT (dg_stage_mfma.hpp): with
  // opF set the stage runs the FP32-contraction form (advection); the FP64
  // blocks opD serve the FP64 matrix-core variant (tiling code 0).
  const float* __restrict__ opF;
  const double* __restrict__ opD;
  // Host-side kernel choice, not read by device code: 1 selects the variant
  // with nontemporal accumulator streams.
  int streaming;
};

using StageFn = void (*)(const StageParams);

// MI355X last-level (Infinity) cache. The RK stage kernels use nontemporal
// accumulator streams when a step's working set exceeds it.
constexpr size_t kLastLevelCacheBytes = size_t(256) << 20;

// Bytes one RK4 step touches: the state and the four work vectors of F
// fields, plus the 4 metric terms and 3 connectivity codes per cell.
inline size_t rk_working_set_bytes(int K, int Np, int F) {
  return size_t(K) * (5 * size_t(F) * Np * sizeof(double) + 4 * sizeof(double) + 3 * sizeof(int));
}

// Tiling code of the FP64 matrix-core variant of the stage kernels.
constexpr int kMfmaTiling = 0;

// Number of operand blocks (16 rows x 4 columns, 64 values each) of the
// matrix-core stage: Dr and Ds as ceil(Np/16) x ceil(Np/4) blocks, LIFT as
// ceil(Np/16) x ceil(3Nfp/4) blocks.
inline int mfma_operand_blocks(int Np, int Nfp) {
  const int MB = (Np + 15) / 16, KS = (Np + 3) / 4, KL = (3 * Nfp + 3) / 4;
  return MB * (2 * KS + KL);
}
hipError_t launch_stage_mfma(int order, int physics, int stage, const StageParams& p, hipStream_t s);

// Batched Maxwell TM ensemble (dg_ensemble.hip): one stage of every instance
// of one order. params: 4 stage parameter records per instance (stages 1..4);
// items: (instance, tile group) pairs of those instances, the items of an
// instance contiguous.
// streaming: the variant with nontemporal accumulator streams.
bool ensemble_supports(int order, int wt);
hipError_t launch_ensemble_stage(int order, int wt, bool streaming, int stage, const StageParams* params,
                                 const int2* items, int nitems, hipStream_t s);

// Padded operator row count used by the stage kernels.
inline int padded_rows(int Np) { return kWaves * ((Np + kWaves - 1) / kWaves); }

// Tilings of the FP64 stage kernels (the `wt` of the entry points). With
// wt = 1, 2 or 4 waves per 64-cell tile a 256-thread workgroup holds 4/wt
// tiles; with wt = 8 or 16 one tile has a 512- or 1024-thread workgroup to
// itself (fewer output rows per wave: a shorter critical path on meshes of a
// few tiles). kSplitTiling spreads one tile over kSplitGroups workgroups of 4
// waves on as many CUs: each loads the tile and forms all its face fluxes,
// and computes a quarter of the output rows, so on a mesh of a few tiles the
// LDS reads and products of a tile are shared by four CUs instead of one.
constexpr int kSplitTiling = 32;
constexpr int kSplitGroups = 4;
constexpr int tile_splits(int wt) { return wt == kSplitTiling ? kSplitGroups : 1; }
// Waves of one workgroup that serve a tile.
constexpr int tile_waves(int wt) { return wt == kSplitTiling ? kWaves : wt; }
constexpr int tiles_per_group(int wt) { return wt >= kWaves ? 1 : kWaves / wt; }
constexpr int group_threads(int wt) { return kTile * (tile_waves(wt) > kWaves ? tile_waves(wt) : kWaves); }

// Maxwell stage kernels where the lift applies the face normal per face: the
// tile keeps two flux values per face node (the upwind combination w / 2
// shared by Hx and Hy, and the Ez flux) instead of three. Measured per order
// and tiling (whole RK4 steps, K = 8.4M at N <= 3, K = 131072 above): N = 7
// at 8 waves per tile 689 -> 513 us, N = 2 at 2 waves 4672 -> 4523 us; at
// N = 3..6 the extra registers cost what the smaller LDS footprint gains.
constexpr bool stage_factors_normal(int N, int physics, int wt) {
  return physics == kMaxwellTM && ((N == 2 && wt <= 2) || N == 7);
}
// Flux values kept per face node.
constexpr int stage_flux_fields(int N, int physics, int wt) {
  return physics != kMaxwellTM ? 1 : (stage_factors_normal(N, physics, wt) ? 2 : 3);
}
// LDS (doubles) of one FP64 stage workgroup: node rows of the F fields and
// face fluxes of its tiles.
constexpr int stage_lds_doubles(int N, int physics, int wt) {
  return tiles_per_group(wt) *
         ((physics == kMaxwellTM ? 3 : 1) * (N + 1) * (N + 2) / 2 + stage_flux_fields(N, physics, wt) * 3 * (N + 1)) *
         kTile;
}

// wt: waves per 64-cell tile (1, 2, 4, 8 or 16) or kSplitTiling; a variant
// whose LDS would not fit falls back to 4 (supported_waves_per_tile tells
// which one runs).
hipError_t launch_stage(int order, int physics, int stage, int wt, const StageParams& p, hipStream_t s);
int supported_waves_per_tile(int order, int physics, int wt);
int default_waves_per_tile(int order, int physics, int K);

}  // namespace dgc

// Error reporting shared by every entry point.
int dgc_set_error(int code, const std::string& msg);
