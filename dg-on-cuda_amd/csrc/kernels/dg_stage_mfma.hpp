// RK4 stage with the volume and lift contractions on the matrix cores, for
// the advection (1 field) and Maxwell TM (3 fields) operators.
//
// Compute type CT = float: the FP32-contraction form (v_mfma_f32_16x16x4f32,
// operands rounded to FP32), both operators. CT = double: a variant of the plain FP64 path
// (v_mfma_f64_16x16x4f64) that the autotuner can pick. The state, the stage
// accumulation, the metric terms, the numerical fluxes and the stage update
// are FP64 in both.
//
// Same contract as the VALU stage (dg_stage_tile.hpp; it replaces
// examples/maxwell_2d/d_maxwell_2d.cuh:50-131,
// examples/advection_2d/d_advection_2d.cuh:45-100 and the axpy passes of
// src/explicit_runge_kutta.h:33-51). One 64-cell tile per 256-thread
// workgroup:
//   1. every HBM read of the tile is issued first (node rows, connectivity,
//      metric terms, neighbour traces outside the tile); for advection the
//      operand blocks are copied to LDS (for Maxwell they are read through L1);
//   2. numerical fluxes at the 3 Nfp face nodes in FP64 (LDS);
//   3. wave w owns cells 16w..16w+15 of the tile as the N-dimension of
//      D[16 x 16] += A[16 x 4] B[4 x 16] products: A = 16-row blocks of Dr,
//      Ds and LIFT, B = node values / fluxes of those cells. The lane holding
//      output row j of cell q combines the products with the cell's metric
//      terms in FP64 and applies the RK4 stage update in FP64.
#pragma once
#include "dg_internal.hpp"
#include "dg_mfma.hpp"

namespace dgc {

namespace mfma_detail {

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

}  // namespace mfma_detail

// Operand blocks are staged in LDS for the single-field operator only; the
// Maxwell tile already holds 3 fields of node rows and fluxes there.
template <int PHYS>
constexpr bool mfma_operands_in_lds() {
  return PHYS != kMaxwellTM;
}

// LDS bytes of one matrix-core stage workgroup.
template <typename CT, int N, int PHYS>
constexpr int mfma_lds_bytes() {
  constexpr int Np = (N + 1) * (N + 2) / 2, NF = 3 * (N + 1), F = PHYS == kMaxwellTM ? 3 : 1;
  constexpr int MB = (Np + 15) / 16, KS = (Np + 3) / 4, KL = (NF + 3) / 4;
  return F * (Np + NF) * kTile * 8 + (mfma_operands_in_lds<PHYS>() ? MB * (2 * KS + KL) * 64 * int(sizeof(CT)) : 0);
}

// NT: nontemporal hints on the accumulator streams (ld_once / st_once).
template <typename CT, int N, int PHYS, int STAGE, bool NT>
__device__ __forceinline__ void stage_tile_mfma(const StageParams& p, int tile, unsigned char* lds) {
  using namespace mfma_detail;
  using T = Tri<N>;
  constexpr int Np = T::Np, Nfp = T::Nfp, NF = 3 * Nfp;
  constexpr int F = PHYS == kMaxwellTM ? 3 : 1;
  constexpr int MB = (Np + 15) / 16, KS = (Np + 3) / 4, KL = (NF + 3) / 4;
  constexpr int NA = MB * (2 * KS + KL);
  constexpr int LPW = (Np + kWaves - 1) / kWaves;
  constexpr int FPW = (NF + kWaves - 1) / kWaves;
  constexpr bool ALDS = mfma_operands_in_lds<PHYS>();
  using UsT = double[Np][kTile];
  using FlT = double[NF][kTile];
  UsT* us = reinterpret_cast<UsT*>(lds);                                   // us[F][Np][kTile]
  FlT* fl = reinterpret_cast<FlT*>(lds + size_t(F) * Np * kTile * 8);      // fl[F][NF][kTile]
  CT* opa = reinterpret_cast<CT*>(lds + size_t(F) * (Np + NF) * kTile * 8);  // [NA][64] if ALDS

  const int K = p.K;
  const int lane = threadIdx.x & (kTile - 1);
  const int wave = uniform(threadIdx.x >> 6);
  const int c0 = tile * kTile;
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
  double tl[F][LPW];
#pragma unroll
  for (int i = 0; i < LPW; ++i) {
    const int j = wave + kWaves * i;
#pragma unroll
    for (int f = 0; f < F; ++f) tl[f][i] = (valid && j < Np) ? p.in[f][(size_t)j * K + c] : 0.0;
  }
  double rx = 0.0, ry = 0.0, sx = 0.0, sy = 0.0;
  if (valid) {
    rx = p.geo[(size_t)kRx * K + c];
    ry = p.geo[(size_t)kRy * K + c];
    sx = p.geo[(size_t)kSx * K + c];
    sy = p.geo[(size_t)kSy * K + c];
  }
  double gv[F][FPW];
#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wave + kWaves * i;
    const int d = m % Nfp;
    const int nf = code[i] & 3, idx = code[i] >> 2;
    const bool far = nf != kBoundaryCode && (unsigned)(idx - c0) >= (unsigned)kTile;
#pragma unroll
    for (int f = 0; f < F; ++f) gv[f][i] = far ? p.in[f][(size_t)T::fnode(nf & 3, Nfp - 1 - d) * K + idx] : 0.0;
  }
  if constexpr (ALDS) {
    const CT* src = operand_blocks<CT>(p);
    for (int i = threadIdx.x; i < NA * 64; i += kBlock) opa[i] = src[i];
  }
#pragma unroll
  for (int i = 0; i < LPW; ++i) {
    const int j = wave + kWaves * i;
    if (j < Np) {
#pragma unroll
      for (int f = 0; f < F; ++f) us[f][j][lane] = tl[f][i];
    }
  }
  __syncthreads();

  // ---- 2. numerical fluxes, scaled by face_J / J --------------------------------
#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wave + kWaves * i;
    if (m >= NF) break;
    const int e = m / Nfp, d = m - e * Nfp;
    const int own = T::fnode(e, d);
    double a[F], b[F];
#pragma unroll
    for (int f = 0; f < F; ++f) a[f] = us[f][own][lane];
    const double vx = e == 0 ? sx : (e == 1 ? rx + sx : rx);
    const double vy = e == 0 ? sy : (e == 1 ? ry + sy : ry);
    const double fs = sqrt(vx * vx + vy * vy);
    const double sg = (e == 1 ? 1.0 : -1.0) / (valid ? fs : 1.0);
    const double nx = sg * vx, ny = sg * vy;
    const int nf = code[i] & 3, idx = code[i] >> 2;
    if (nf == kBoundaryCode) {
      if constexpr (PHYS == kMaxwellTM) {
        // PEC wall: (Hx, Hy, Ez)+ = (Hx, Hy, -Ez)-
        b[0] = a[0];
        b[1 % F] = a[1 % F];
        b[2 % F] = -a[2 % F];
      } else if (p.bc == kBcSinSin && valid) {
        const double xb = p.bx[idx + d], yb = p.by[idx + d];
        b[0] = sin((xb + 1.0 - t) * M_PI) * sin((yb + 1.0 - t) * M_PI);
      } else {
        b[0] = 0.0;
      }
    } else {
      const int l = idx - c0;
      const int nn = T::fnode(nf, Nfp - 1 - d);
#pragma unroll
      for (int f = 0; f < F; ++f) b[f] = (unsigned)l < (unsigned)kTile ? us[f][nn][l] : gv[f][i];
    }
    if constexpr (PHYS == kMaxwellTM) {
      const double dHx = a[0] - b[0];
      const double dHy = a[1 % F] - b[1 % F];
      const double sEz = a[2 % F] + b[2 % F];
      const double w = sEz + ny * dHx - nx * dHy;
      fl[0][m][lane] = 0.5 * ny * w * fs;
      fl[1 % F][m][lane] = -0.5 * nx * w * fs;
      fl[2 % F][m][lane] = 0.5 * (ny * (a[0] + b[0]) - nx * (a[1 % F] + b[1 % F]) + a[2 % F] - b[2 % F]) * fs;
    } else {
      const double an = p.ax * nx + p.ay * ny;
      fl[0][m][lane] = (an >= 0.0 ? a[0] : b[0]) * an * fs;
    }
  }
  __syncthreads();

  // ---- 3. contractions on the matrix cores -------------------------------------
  const int col = lane & 15, kq = lane >> 4;
  const int q = 16 * wave + col;  // tile cell of this lane's output column
  const int cc = c0 + q;
  const bool ovalid = cc < K;
  double qrx = 0.0, qry = 0.0, qsx = 0.0, qsy = 0.0;
  if (ovalid) {
    qrx = p.geo[(size_t)kRx * K + cc];
    qry = p.geo[(size_t)kRy * K + cc];
    qsx = p.geo[(size_t)kSx * K + cc];
    qsy = p.geo[(size_t)kSy * K + cc];
  }
  const CT* A = ALDS ? opa : operand_blocks<CT>(p);
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) {
    // stage operands of this lane's four rows, in flight during the products
    double o1[F][4], o2[F][4], o3[F][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = 16 * mb + Acc<CT>::row(kq, r);
      const size_t idx = (size_t)j * K + cc;
#pragma unroll
      for (int f = 0; f < F; ++f) {
        o1[f][r] = o2[f][r] = o3[f][r] = 0.0;
        if (ovalid && j < Np) {
          if (STAGE == 2) {
            o1[f][r] = ld_once<NT>(p.acc[f] + idx);
            o2[f][r] = ld_once<NT>(p.u[f] + idx);
          } else if (STAGE == 3) {
            o2[f][r] = ld_once<NT>(p.u[f] + idx);
          } else if (STAGE == 4) {
            o1[f][r] = ld_once<NT>(p.acc[f] + idx);
            o2[f][r] = ld_once<NT>(p.u[f] + idx);
            o3[f][r] = ld_once<NT>(p.k3[f] + idx);
          }
        }
      }
    }
    typename Acc<CT>::type ar[F], as[F], al[F];
#pragma unroll
    for (int f = 0; f < F; ++f) ar[f] = as[f] = al[f] = typename Acc<CT>::type{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const CT a_r = A[(mb * KS + ks) * 64 + lane];
      const CT a_s = A[(MB * KS + mb * KS + ks) * 64 + lane];
      const int k = 4 * ks + kq;
#pragma unroll
      for (int f = 0; f < F; ++f) {
        const CT bv = k < Np ? CT(us[f][k][q]) : CT(0);
        ar[f] = mma(a_r, bv, ar[f]);
        as[f] = mma(a_s, bv, as[f]);
      }
    }
#pragma unroll
    for (int ks = 0; ks < KL; ++ks) {
      const CT a_l = A[(2 * MB * KS + mb * KL + ks) * 64 + lane];
      const int m = 4 * ks + kq;
#pragma unroll
      for (int f = 0; f < F; ++f) al[f] = mma(a_l, m < NF ? CT(fl[f][m][q]) : CT(0), al[f]);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = 16 * mb + Acc<CT>::row(kq, r);
      if (!ovalid || j >= Np) continue;
      double res[F];
      if constexpr (PHYS == kMaxwellTM) {
        // Hx_t = Dy Ez, Hy_t = -Dx Ez, Ez_t = Dy Hx - Dx Hy, then minus the lift
        res[0] = qry * double(ar[2 % F][r]) + qsy * double(as[2 % F][r]) - double(al[0][r]);
        res[1 % F] = -(qrx * double(ar[2 % F][r]) + qsx * double(as[2 % F][r])) - double(al[1 % F][r]);
        res[2 % F] = (qry * double(ar[0][r]) + qsy * double(as[0][r])) -
                     (qrx * double(ar[1 % F][r]) + qsx * double(as[1 % F][r])) - double(al[2 % F][r]);
      } else {
        const double alpha = p.ax * qrx + p.ay * qry, beta = p.ax * qsx + p.ay * qsy;
        res[0] = alpha * double(ar[0][r]) + beta * double(as[0][r]) - double(al[0][r]);
      }
      const size_t idx = (size_t)j * K + cc;
#pragma unroll
      for (int f = 0; f < F; ++f) {
        if (STAGE == 0) {
          p.out[f][idx] = res[f];
        } else if (STAGE == 1) {
          st_once<NT>(p.acc[f] + idx, res[f]);
          p.out[f][idx] = p.c1 * res[f] + us[f][j][q];
        } else if (STAGE == 2) {
          st_once<NT>(p.acc[f] + idx, 2.0 * res[f] + o1[f][r]);
          p.out[f][idx] = p.c1 * res[f] + o2[f][r];
        } else if (STAGE == 3) {
          st_once<NT>(p.k3[f] + idx, res[f]);
          p.out[f][idx] = p.c1 * res[f] + o2[f][r];
        } else {
          const double w1 = 2.0 * o3[f][r] + res[f];
          const double w2 = p.c1 * o1[f][r] + o2[f][r];
          st_once<NT>(p.u[f] + idx, p.c1 * w1 + w2);
        }
      }
    }
  }
}

template <typename CT, int N, int PHYS, int STAGE, bool NT>
__global__ __launch_bounds__(kBlock) void dg_stage_mfma_kernel(const StageParams p) {
  __shared__ __align__(16) unsigned char lds[mfma_lds_bytes<CT, N, PHYS>()];
  stage_tile_mfma<CT, N, PHYS, STAGE, NT>(p, xcd_group(blockIdx.x, gridDim.x), lds);
}

namespace mfma_detail {

template <typename CT, int N, int PHYS, bool NT>
StageFn stage_fn(int stage) {
  return stage == 0 ? dg_stage_mfma_kernel<CT, N, PHYS, 0, NT>
       : stage == 1 ? dg_stage_mfma_kernel<CT, N, PHYS, 1, NT>
       : stage == 2 ? dg_stage_mfma_kernel<CT, N, PHYS, 2, NT>
       : stage == 3 ? dg_stage_mfma_kernel<CT, N, PHYS, 3, NT>
                    : dg_stage_mfma_kernel<CT, N, PHYS, 4, NT>;
}

template <typename CT, int PHYS, bool NT>
StageFn pick_order(int order, int stage) {
  switch (order) {
    case 1: return stage_fn<CT, 1, PHYS, NT>(stage);
    case 2: return stage_fn<CT, 2, PHYS, NT>(stage);
    case 3: return stage_fn<CT, 3, PHYS, NT>(stage);
    case 4: return stage_fn<CT, 4, PHYS, NT>(stage);
    case 5: return stage_fn<CT, 5, PHYS, NT>(stage);
    case 6: return stage_fn<CT, 6, PHYS, NT>(stage);
    case 7: return stage_fn<CT, 7, PHYS, NT>(stage);
    default: return nullptr;
  }
}

// Kernel of the FP32-contraction form (f32) or of the FP64 matrix-core
// variant, for either operator; nullptr if not built.
template <bool NT>
StageFn pick(int order, int physics, bool f32, int stage) {
  if (f32)
    return physics == kMaxwellTM ? pick_order<float, kMaxwellTM, NT>(order, stage)
                                 : pick_order<float, kAdvection, NT>(order, stage);
  return physics == kMaxwellTM ? pick_order<double, kMaxwellTM, NT>(order, stage)
                               : pick_order<double, kAdvection, NT>(order, stage);
}

// The nontemporal table (dg_rk_mfma_nt.hip).
StageFn pick_streaming(int order, int physics, bool f32, int stage);

}  // namespace mfma_detail

}  // namespace dgc
