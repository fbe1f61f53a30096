// Fused nodal-DG right-hand side + classical RK4 stage update (FP64).
//
// Replaces, as one kernel per stage:
//   examples/maxwell_2d/d_maxwell_2d.cuh:50-131   (Maxwell TM volume, upwind/PEC flux, LIFT)
//   examples/advection_2d/d_advection_2d.cuh:45-100 (advection volume, upwind flux, LIFT)
//   src/device_SemiDiscOp_wrapper.cuh:35-58        (grid-stride launcher)
//   src/explicit_runge_kutta.h:33-51 + src/axpy.h  (6 axpy passes per step, fused here)
//
// Mapping: one workgroup = 4 waves = a tile of 64 consecutive cells (lane =
// cell). The tile's node rows are staged in LDS with coalesced 512-byte row
// loads; the face traces of neighbours that live in the same tile are read
// from LDS, others from global memory (L2). The face fluxes are computed
// once per (cell, face node) and kept in LDS. Each wave then owns a block of
// ceil(Np/4) output rows: the Dr/Ds/LIFT coefficients of those rows are
// wave-uniform, so they are fetched with scalar loads into SGPRs and every
// LDS read of a DOF feeds 2*rows FMAs. The RK4 stage update is applied
// while the result is in registers, so each stage moves the state once.
#include "dg_internal.hpp"

namespace dgc {

template <int N, int PHYS, int STAGE>
__global__ __launch_bounds__(kBlock) void dg_stage_kernel(const StageParams p) {
  using T = Tri<N>;
  constexpr int Np = T::Np, Nfp = T::Nfp, NF = 3 * Nfp;
  constexpr int F = PHYS == kMaxwellTM ? 3 : 1;
  constexpr int RPW = (Np + kWaves - 1) / kWaves;
  constexpr int Npad = RPW * kWaves;
  __shared__ double us[F][Np][kTile];
  __shared__ double fl[F][NF][kTile];

  const int K = p.K;
  const int lane = threadIdx.x & (kTile - 1);
  const int wave = uniform(threadIdx.x >> 6);
  const int c0 = blockIdx.x * kTile;
  const int c = c0 + lane;
  const bool valid = c < K;

  double t = p.t;
  if (p.tclock != nullptr) t = p.tclock[0] + p.tfrac * p.dt_clock;

  // 1. stage the tile's node rows in LDS
  for (int j = wave; j < Np; j += kWaves) {
#pragma unroll
    for (int f = 0; f < F; ++f) us[f][j][lane] = valid ? p.in[f][(size_t)j * K + c] : 0.0;
  }
  __syncthreads();

  // 2. numerical fluxes at the face nodes, pre-scaled by face_J / J
  for (int m = wave; m < NF; m += kWaves) {
    const int e = m / Nfp;
    const int d = m - e * Nfp;
    const int own = T::fnode(e, d);
    double a[F], b[F];
#pragma unroll
    for (int f = 0; f < F; ++f) a[f] = us[f][own][lane];
    double nx = 0.0, ny = 0.0, fs = 0.0;
    bool boundary = false;
    if (valid) {
      nx = p.geo[(size_t)(kNx0 + e) * K + c];
      ny = p.geo[(size_t)(kNy0 + e) * K + c];
      fs = p.geo[(size_t)(kFs0 + e) * K + c];
      const int code = p.nbr[(size_t)e * K + c];
      const int nf = code & 3;
      const int idx = code >> 2;
      if (nf == kBoundaryCode) {
        boundary = true;
        if (PHYS == kMaxwellTM) {
          // PEC wall: (Hx, Hy, Ez)+ = (Hx, Hy, -Ez)-
          b[0] = a[0];
          if (F == 3) {
            b[1] = a[1];
            b[2] = -a[2];
          }
        } else {
          if (p.bc == kBcSinSin) {
            const double xb = p.bx[idx + d], yb = p.by[idx + d];
            b[0] = sin((xb + 1.0 - t) * M_PI) * sin((yb + 1.0 - t) * M_PI);
          } else {
            b[0] = 0.0;
          }
        }
      } else {
        // neighbour traces run against this face's orientation
        const int nn = T::fnode(nf, Nfp - 1 - d);
        const int l = idx - c0;
        if ((unsigned)l < (unsigned)kTile) {
#pragma unroll
          for (int f = 0; f < F; ++f) b[f] = us[f][nn][l];
        } else {
#pragma unroll
          for (int f = 0; f < F; ++f) b[f] = p.in[f][(size_t)nn * K + idx];
        }
      }
    } else {
#pragma unroll
      for (int f = 0; f < F; ++f) b[f] = a[f];
    }
    (void)boundary;
    if (PHYS == kMaxwellTM) {
      const double dHx = a[0] - b[0];
      const double dHy = a[1 % F] - b[1 % F];
      const double sEz = a[2 % F] + b[2 % F];
      const double w = sEz + ny * dHx - nx * dHy;
      fl[0][m][lane] = 0.5 * ny * w * fs;
      if (F == 3) {
        fl[1 % F][m][lane] = -0.5 * nx * w * fs;
        fl[2 % F][m][lane] =
            0.5 * (ny * (a[0] + b[0]) - nx * (a[1 % F] + b[1 % F]) + a[2 % F] - b[2 % F]) * fs;
      }
    } else {
      const double an = p.ax * nx + p.ay * ny;
      const double U = an >= 0.0 ? a[0] : b[0];
      fl[0][m][lane] = U * an * fs;
    }
  }
  __syncthreads();

  // 3. volume term for this wave's rows: Dr.u and Ds.u per field
  const int j0 = wave * RPW;
  double rr[F][RPW], ss[F][RPW];
#pragma unroll
  for (int f = 0; f < F; ++f)
#pragma unroll
    for (int i = 0; i < RPW; ++i) rr[f][i] = ss[f][i] = 0.0;
#pragma unroll
  for (int k = 0; k < Np; ++k) {
    double uk[F];
#pragma unroll
    for (int f = 0; f < F; ++f) uk[f] = us[f][k][lane];
    const double* dr = p.DrT + k * Npad + j0;
    const double* ds = p.DsT + k * Npad + j0;
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      const double cr = dr[i], cs = ds[i];
#pragma unroll
      for (int f = 0; f < F; ++f) {
        rr[f][i] = fma(cr, uk[f], rr[f][i]);
        ss[f][i] = fma(cs, uk[f], ss[f][i]);
      }
    }
  }

  double res[F][RPW];
  {
    double rx = 0.0, ry = 0.0, sx = 0.0, sy = 0.0;
    if (valid) {
      rx = p.geo[(size_t)kRx * K + c];
      ry = p.geo[(size_t)kRy * K + c];
      sx = p.geo[(size_t)kSx * K + c];
      sy = p.geo[(size_t)kSy * K + c];
    }
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      if (PHYS == kMaxwellTM) {
        // Hx_t = Dy Ez, Hy_t = -Dx Ez, Ez_t = Dy Hx - Dx Hy
        res[0][i] = ry * rr[2 % F][i] + sy * ss[2 % F][i];
        if (F == 3) {
          res[1 % F][i] = -(rx * rr[2 % F][i] + sx * ss[2 % F][i]);
          res[2 % F][i] = (ry * rr[0][i] + sy * ss[0][i]) - (rx * rr[1 % F][i] + sx * ss[1 % F][i]);
        }
      } else {
        res[0][i] = (p.ax * rx + p.ay * ry) * rr[0][i] + (p.ax * sx + p.ay * sy) * ss[0][i];
      }
    }
  }

  // 4. lift: res -= LIFT . (flux * face_J / J)
#pragma unroll
  for (int m = 0; m < NF; ++m) {
    double fm[F];
#pragma unroll
    for (int f = 0; f < F; ++f) fm[f] = fl[f][m][lane];
    const double* lt = p.LT + m * Npad + j0;
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      const double cl = lt[i];
#pragma unroll
      for (int f = 0; f < F; ++f) res[f][i] = fma(-cl, fm[f], res[f][i]);
    }
  }

  // 5. stage update (reference operand order: out = a * x + y)
  if (!valid) return;
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    const int j = j0 + i;
    if (j >= Np) break;
    const size_t idx = (size_t)j * K + c;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const double r = res[f][i];
      if (STAGE == 0) {
        p.out[f][idx] = r;
      } else if (STAGE == 1) {
        p.acc[f][idx] = r;
        p.out[f][idx] = p.c1 * r + us[f][j][lane];
      } else if (STAGE == 2) {
        p.acc[f][idx] = 2.0 * r + p.acc[f][idx];
        p.out[f][idx] = p.c1 * r + p.u[f][idx];
      } else if (STAGE == 3) {
        p.k3[f][idx] = r;
        p.out[f][idx] = p.c1 * r + p.u[f][idx];
      } else {
        const double w1 = 2.0 * p.k3[f][idx] + r;
        const double w2 = p.c1 * p.acc[f][idx] + p.u[f][idx];
        p.u[f][idx] = p.c1 * w1 + w2;
      }
    }
  }
}

namespace {

using StageFn = void (*)(const StageParams);

template <int N, int PHYS>
constexpr StageFn stage_fn(int stage) {
  return stage == 0 ? dg_stage_kernel<N, PHYS, 0>
       : stage == 1 ? dg_stage_kernel<N, PHYS, 1>
       : stage == 2 ? dg_stage_kernel<N, PHYS, 2>
       : stage == 3 ? dg_stage_kernel<N, PHYS, 3>
                    : dg_stage_kernel<N, PHYS, 4>;
}

template <int PHYS>
StageFn pick(int order, int stage) {
  switch (order) {
    case 1: return stage_fn<1, PHYS>(stage);
    case 2: return stage_fn<2, PHYS>(stage);
    case 3: return stage_fn<3, PHYS>(stage);
    case 4: return stage_fn<4, PHYS>(stage);
    case 5: return stage_fn<5, PHYS>(stage);
    case 6: return stage_fn<6, PHYS>(stage);
    case 7: return stage_fn<7, PHYS>(stage);
    default: return nullptr;
  }
}

}  // namespace

hipError_t launch_stage(int order, int physics, int stage, const StageParams& p, hipStream_t s) {
  if (stage < 0 || stage > 4) return hipErrorInvalidValue;
  StageFn fn = physics == kMaxwellTM ? pick<kMaxwellTM>(order, stage) : pick<kAdvection>(order, stage);
  if (fn == nullptr) return hipErrorInvalidValue;
  const int grid = (p.K + kTile - 1) / kTile;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(kBlock), 0, s, p);
  return hipGetLastError();
}

}  // namespace dgc
