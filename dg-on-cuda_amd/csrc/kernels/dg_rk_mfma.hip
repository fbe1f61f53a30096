// Advection RK4 stage with the volume and lift contractions on the matrix
// cores, in FP32 (v_mfma_f32_16x16x4f32: the FP32-contraction form) or in
// FP64 (v_mfma_f64_16x16x4f64: a variant of the plain FP64 path that the
// autotuner can pick). The state, the stage accumulation, the metric terms,
// the numerical fluxes and the stage update are FP64 in both.
//
// Same contract as the FP64 stage kernel (dg_rk.hip; it replaces
// examples/advection_2d/d_advection_2d.cuh:45-100 and the axpy passes of
// src/explicit_runge_kutta.h:33-51). Per 64-cell tile and 256-thread
// workgroup:
//   1. every HBM read of the tile is issued first (node rows, connectivity,
//      metric terms, neighbour traces outside the tile); the operand blocks
//      are copied to LDS;
//   2. upwind fluxes at the 3 Nfp face nodes in FP64 (LDS);
//   3. wave w owns cells 16w..16w+15 of the tile as the N-dimension of
//      D[16 x 16] += A[16 x 4] B[4 x 16] products: A = 16-row blocks of Dr,
//      Ds and LIFT, B = node values / fluxes (rounded to FP32 in the FP32
//      form). The lane holding output row j of cell q combines
//         (a_x r_x + a_y r_y)(Dr u)_j + (a_x s_x + a_y s_y)(Ds u)_j - (LIFT f)_j
//      in FP64 and applies the RK4 stage update in FP64.
#include "dg_internal.hpp"
#include "dg_mfma.hpp"

namespace dgc {

namespace {

// Workgroup renumbering so that each of the 8 XCDs sweeps a contiguous range
// of tiles (same mapping as the FP64 stage kernel).
__device__ __forceinline__ int xcd_tile(int b, int G) {
  constexpr int X = 8;
  const int x = b % X, i = b / X, per = G / X, rem = G % X;
  return x * per + (x < rem ? x : rem) + i;
}

__device__ __forceinline__ floatx4 mma(float a, float b, floatx4 c) { return mfma16x16x4f32(a, b, c); }
__device__ __forceinline__ doublex4 mma(double a, double b, doublex4 c) { return mfma16x16x4f64(a, b, c); }

// Result type and row of result register r in lane group kq = lane / 16:
// fp32 rows are blocked (4 kq + r), fp64 rows interleaved (4 r + kq).
template <typename CT>
struct Acc;
template <>
struct Acc<float> {
  using type = floatx4;
  static constexpr __device__ int row(int kq, int r) { return 4 * kq + r; }
};
template <>
struct Acc<double> {
  using type = doublex4;
  static constexpr __device__ int row(int kq, int r) { return 4 * r + kq; }
};

template <typename CT>
__device__ __forceinline__ const CT* operand_blocks(const StageParams& p);
template <>
__device__ __forceinline__ const float* operand_blocks<float>(const StageParams& p) {
  return p.opF;
}
template <>
__device__ __forceinline__ const double* operand_blocks<double>(const StageParams& p) {
  return p.opD;
}

template <typename CT, int N, int STAGE>
__global__ __launch_bounds__(kBlock) void dg_stage_mfma_kernel(const StageParams p) {
  using T = Tri<N>;
  constexpr int Np = T::Np, Nfp = T::Nfp, NF = 3 * Nfp;
  constexpr int MB = (Np + 15) / 16, KS = (Np + 3) / 4, KL = (NF + 3) / 4;
  constexpr int NA = MB * (2 * KS + KL);
  constexpr int LPW = (Np + kWaves - 1) / kWaves;
  constexpr int FPW = (NF + kWaves - 1) / kWaves;
  __shared__ double us[Np][kTile];
  __shared__ double fl[NF][kTile];
  __shared__ CT opa[NA][64];

  const int K = p.K;
  const int lane = threadIdx.x & (kTile - 1);
  const int wave = uniform(threadIdx.x >> 6);
  const int c0 = xcd_tile(blockIdx.x, gridDim.x) * kTile;
  const int c = c0 + lane;
  const bool valid = c < K;
  double t = p.t;
  if (p.tclock != nullptr) t = p.tclock[0] + p.tfrac * p.dt_clock;

  // ---- 1. loads --------------------------------------------------------------
  int code[FPW];
#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wave + kWaves * i;
    code[i] = (valid && m < NF) ? p.nbr[(size_t)(m / Nfp) * K + c] : kBoundaryCode;
  }
  double tl[LPW];
#pragma unroll
  for (int i = 0; i < LPW; ++i) {
    const int j = wave + kWaves * i;
    tl[i] = (valid && j < Np) ? p.in[0][(size_t)j * K + c] : 0.0;
  }
  double rx = 0.0, ry = 0.0, sx = 0.0, sy = 0.0;
  if (valid) {
    rx = p.geo[(size_t)kRx * K + c];
    ry = p.geo[(size_t)kRy * K + c];
    sx = p.geo[(size_t)kSx * K + c];
    sy = p.geo[(size_t)kSy * K + c];
  }
  double gv[FPW];
#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wave + kWaves * i;
    const int d = m % Nfp;
    const int nf = code[i] & 3, idx = code[i] >> 2;
    const bool far = nf != kBoundaryCode && (unsigned)(idx - c0) >= (unsigned)kTile;
    gv[i] = far ? p.in[0][(size_t)T::fnode(nf & 3, Nfp - 1 - d) * K + idx] : 0.0;
  }
  {
    const CT* src = operand_blocks<CT>(p);
    for (int i = threadIdx.x; i < NA * 64; i += kBlock) (&opa[0][0])[i] = src[i];
  }
#pragma unroll
  for (int i = 0; i < LPW; ++i) {
    const int j = wave + kWaves * i;
    if (j < Np) us[j][lane] = tl[i];
  }
  __syncthreads();

  // ---- 2. upwind fluxes, scaled by face_J / J ----------------------------------
#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wave + kWaves * i;
    if (m >= NF) break;
    const int e = m / Nfp, d = m - e * Nfp;
    const double a = us[T::fnode(e, d)][lane];
    const double vx = e == 0 ? sx : (e == 1 ? rx + sx : rx);
    const double vy = e == 0 ? sy : (e == 1 ? ry + sy : ry);
    const double fs = sqrt(vx * vx + vy * vy);
    const double sg = (e == 1 ? 1.0 : -1.0) / (valid ? fs : 1.0);
    const double nx = sg * vx, ny = sg * vy;
    const int nf = code[i] & 3, idx = code[i] >> 2;
    double b;
    if (nf == kBoundaryCode) {
      if (p.bc == kBcSinSin && valid) {
        const double xb = p.bx[idx + d], yb = p.by[idx + d];
        b = sin((xb + 1.0 - t) * M_PI) * sin((yb + 1.0 - t) * M_PI);
      } else {
        b = 0.0;
      }
    } else {
      const int l = idx - c0;
      b = (unsigned)l < (unsigned)kTile ? us[T::fnode(nf, Nfp - 1 - d)][l] : gv[i];
    }
    const double an = p.ax * nx + p.ay * ny;
    fl[m][lane] = (an >= 0.0 ? a : b) * an * fs;
  }
  __syncthreads();

  // ---- 3. contractions on the matrix cores -------------------------------------
  const int col = lane & 15, kq = lane >> 4;
  const int q = 16 * wave + col;  // tile cell of this lane's output column
  const int cc = c0 + q;
  const bool ovalid = cc < K;
  double alpha = 0.0, beta = 0.0;
  if (ovalid) {
    alpha = p.ax * p.geo[(size_t)kRx * K + cc] + p.ay * p.geo[(size_t)kRy * K + cc];
    beta = p.ax * p.geo[(size_t)kSx * K + cc] + p.ay * p.geo[(size_t)kSy * K + cc];
  }
  CT bu[KS], bl[KL];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int k = 4 * ks + kq;
    bu[ks] = k < Np ? CT(us[k][q]) : CT(0);
  }
#pragma unroll
  for (int ks = 0; ks < KL; ++ks) {
    const int m = 4 * ks + kq;
    bl[ks] = m < NF ? CT(fl[m][q]) : CT(0);
  }
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) {
    // stage operands of this lane's four rows, in flight during the products
    double o1[4], o2[4], o3[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = 16 * mb + Acc<CT>::row(kq, r);
      const size_t idx = (size_t)j * K + cc;
      o1[r] = o2[r] = o3[r] = 0.0;
      if (ovalid && j < Np) {
        if (STAGE == 2) {
          o1[r] = p.acc[0][idx];
          o2[r] = p.u[0][idx];
        } else if (STAGE == 3) {
          o2[r] = p.u[0][idx];
        } else if (STAGE == 4) {
          o1[r] = p.acc[0][idx];
          o2[r] = p.u[0][idx];
          o3[r] = p.k3[0][idx];
        }
      }
    }
    typename Acc<CT>::type ar = {0, 0, 0, 0}, as = ar, al = ar;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      ar = mma(opa[mb * KS + ks][lane], bu[ks], ar);
      as = mma(opa[MB * KS + mb * KS + ks][lane], bu[ks], as);
    }
#pragma unroll
    for (int ks = 0; ks < KL; ++ks) al = mma(opa[2 * MB * KS + mb * KL + ks][lane], bl[ks], al);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = 16 * mb + Acc<CT>::row(kq, r);
      if (!ovalid || j >= Np) continue;
      const double res = alpha * double(ar[r]) + beta * double(as[r]) - double(al[r]);
      const size_t idx = (size_t)j * K + cc;
      if (STAGE == 0) {
        p.out[0][idx] = res;
      } else if (STAGE == 1) {
        p.acc[0][idx] = res;
        p.out[0][idx] = p.c1 * res + us[j][q];
      } else if (STAGE == 2) {
        p.acc[0][idx] = 2.0 * res + o1[r];
        p.out[0][idx] = p.c1 * res + o2[r];
      } else if (STAGE == 3) {
        p.k3[0][idx] = res;
        p.out[0][idx] = p.c1 * res + o2[r];
      } else {
        const double w1 = 2.0 * o3[r] + res;
        const double w2 = p.c1 * o1[r] + o2[r];
        p.u[0][idx] = p.c1 * w1 + w2;
      }
    }
  }
}

using StageFn = void (*)(const StageParams);

template <typename CT, int N>
StageFn mfma_fn(int stage) {
  return stage == 0 ? dg_stage_mfma_kernel<CT, N, 0>
       : stage == 1 ? dg_stage_mfma_kernel<CT, N, 1>
       : stage == 2 ? dg_stage_mfma_kernel<CT, N, 2>
       : stage == 3 ? dg_stage_mfma_kernel<CT, N, 3>
                    : dg_stage_mfma_kernel<CT, N, 4>;
}

template <typename CT>
StageFn pick_mfma(int order, int stage) {
  switch (order) {
    case 1: return mfma_fn<CT, 1>(stage);
    case 2: return mfma_fn<CT, 2>(stage);
    case 3: return mfma_fn<CT, 3>(stage);
    case 4: return mfma_fn<CT, 4>(stage);
    case 5: return mfma_fn<CT, 5>(stage);
    case 6: return mfma_fn<CT, 6>(stage);
    case 7: return mfma_fn<CT, 7>(stage);
    default: return nullptr;
  }
}

}  // namespace

hipError_t launch_stage_mfma(int order, int stage, const StageParams& p, hipStream_t s) {
  if (stage < 0 || stage > 4) return hipErrorInvalidValue;
  // the FP32 blocks select the FP32-contraction form; else the FP64 variant
  StageFn fn = p.opF != nullptr ? pick_mfma<float>(order, stage)
             : p.opD != nullptr ? pick_mfma<double>(order, stage) : nullptr;
  if (fn == nullptr) return hipErrorInvalidValue;
  const int grid = (p.K + kTile - 1) / kTile;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(kBlock), 0, s, p);
  return hipGetLastError();
}

}  // namespace dgc
