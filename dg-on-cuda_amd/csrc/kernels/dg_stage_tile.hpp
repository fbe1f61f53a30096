// Fused nodal-DG right-hand side + classical RK4 stage update (FP64).
//
// Replaces, as one kernel per stage:
//   examples/maxwell_2d/d_maxwell_2d.cuh:50-131   (Maxwell TM volume, upwind/PEC flux, LIFT)
//   examples/advection_2d/d_advection_2d.cuh:45-100 (advection volume, upwind flux, LIFT)
//   src/device_SemiDiscOp_wrapper.cuh:35-58        (grid-stride launcher)
//   src/explicit_runge_kutta.h:33-51 + src/axpy.h  (6 axpy passes per step, fused here)
//
// Mapping. A tile is 64 consecutive cells (lane = cell) served by WT waves
// (WT = 1, 2 or 4, chosen per order by the autotuner, and 8 or 16 for meshes
// of a few tiles); a 256-thread workgroup holds 4/WT tiles (WT = 8, 16: one
// tile per 512- or 1024-thread workgroup), and workgroups are renumbered so that each XCD
// (own L2) sweeps a contiguous range of tiles, keeping neighbour traces in
// the L2 that loaded them.
//
// Memory. Everything a tile needs from HBM is requested before the first
// barrier: the connectivity codes, the node rows of the stage input, the four
// metric terms (rx, ry, sx, sy) and the neighbour traces that live outside
// the tile. Face normals and face_J/J are derived from the metric terms
// (straight-sided cells), so a stage reads 4 doubles + 3 ints of geometry per
// cell instead of 13 doubles. Node rows go to LDS; neighbours inside the tile
// are read from LDS.
//
// Compute. Each wave owns ceil(Np/WT) output rows: their Dr/Ds/LIFT
// coefficients are wave-uniform, so they arrive through scalar loads, and
// each LDS read of a DOF feeds 2*rows FMAs. The RK4 stage update is applied
// while the result is in registers (operand order of the reference's axpy).
#pragma once
#include "dg_internal.hpp"

namespace dgc {

// Code shape of the volume and lift loops, measured per operator, order and
// tiling (whole RK4 steps at K = 8.4M for N <= 3 and K = 131072 above).
// Unrolled completely, the loops let the compiler hoist every LDS read and
// coefficient load, which for three fields costs 190-256 VGPRs (1-2 waves per
// SIMD); unrolled by 4 with the stage operands loaded after the products, the
// Maxwell kernels hold 66-110 VGPRs (Maxwell N = 6 at 8 waves per tile: 907
// -> 4282 GB/s; N = 3 at 4 waves: 3601 -> 4495 GB/s). The single-field
// advection kernels keep the complete unroll, except at 8 waves per tile for
// N >= 6 (1319 -> 3568 GB/s at N = 6). At 16 waves per tile and N <= 4 a wave
// owns one output row, so the complete unroll costs few registers and issues
// every coefficient load in one batch, which on a mesh of a few tiles (where
// each launch is a chain of dependent latencies) is what counts: the Maxwell
// cavity of 8 tiles at N = 4, 491 steps, 14.06 -> 13.02 ms.
constexpr int stage_unroll(int N, int PHYS, int WT) {
  if (PHYS == kMaxwellTM) return ((N == 2 && WT <= 2) || (WT >= 16 && N <= 4)) ? 64 : 4;
  return (WT == 8 && N >= 6) ? 2 : 64;
}
constexpr bool stage_delays_operands(int N, int PHYS, int WT) { return stage_unroll(N, PHYS, WT) < 64; }

// One workgroup's share of a stage: the TPW = 4/WT tiles of tile group
// `group` (cells group*64*TPW ...), with `lds` holding stage_lds_doubles()
// doubles; with the split tiling, part `group % RS` of the output rows of tile
// `group / RS`. Used by the per-order kernels below and by the ensemble kernel.
// NT: nontemporal hints on the accumulator streams (ld_once / st_once).
template <int N, int PHYS, int STAGE, int WT, bool NT>
__device__ __forceinline__ void stage_tile(const StageParams& p, int group, double* lds) {
  using T = Tri<N>;
  constexpr int Np = T::Np, Nfp = T::Nfp, NF = 3 * Nfp;
  constexpr int F = PHYS == kMaxwellTM ? 3 : 1;
  constexpr int TPW = tiles_per_group(WT);  // tiles per workgroup
  constexpr int WL = tile_waves(WT);        // waves of this workgroup per tile
  constexpr int RS = tile_splits(WT);       // workgroups sharing a tile's output rows
  constexpr int RPW = (Np + WL * RS - 1) / (WL * RS);  // output rows per wave (contiguous block)
  constexpr int LPW = (Np + WL - 1) / WL;  // input rows loaded per wave (strided)
  constexpr int FPW = (NF + WL - 1) / WL;  // face slots per wave (strided)
  constexpr int S = kWaves * ((Np + kWaves - 1) / kWaves);  // operator row stride
  constexpr bool kDelayOperands = stage_delays_operands(N, PHYS, WT);
  constexpr int kUnroll = stage_unroll(N, PHYS, WT);
  constexpr bool kFactor = stage_factors_normal(N, PHYS, WT);
  constexpr int FL = stage_flux_fields(N, PHYS, WT);
  using UsT = double[F][Np][kTile];
  using FlT = double[FL][NF][kTile];
  UsT* us = reinterpret_cast<UsT*>(lds);                        // us[TPW][F][Np][kTile]
  FlT* fl = reinterpret_cast<FlT*>(lds + TPW * F * Np * kTile);  // fl[TPW][FL][NF][kTile]

  const int K = p.K;
  const int lane = threadIdx.x & (kTile - 1);
  const int wave = uniform(threadIdx.x >> 6);
  const int tw = wave / WL;
  const int wt = wave - tw * WL;
  const int part = RS > 1 ? group % RS : 0;  // output-row part of the split tiling
  const int c0 = ((RS > 1 ? group / RS : group) * TPW + tw) * kTile;
  const int c = c0 + lane;
  const bool valid = c < K;

  double t = p.t;
  if (p.tclock != nullptr) t = p.tclock[0] + p.tfrac * p.dt_clock;

  // ---- 1. issue every independent HBM read of the tile --------------------
  int code[FPW];
#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wt + WL * i;
    code[i] = (valid && m < NF) ? p.nbr[(size_t)(m / Nfp) * K + c] : kBoundaryCode;
  }
  double tl[F][LPW];
#pragma unroll
  for (int i = 0; i < LPW; ++i) {
    const int j = wt + WL * i;
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
  double gv[F][FPW];  // neighbour traces outside the tile
#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wt + WL * i;
    const int e = m / Nfp, d = m - e * Nfp;
    const int nf = code[i] & 3, idx = code[i] >> 2;
    const bool far = nf != kBoundaryCode && (unsigned)(idx - c0) >= (unsigned)kTile;
    const int nn = T::fnode(nf & 3, Nfp - 1 - d);
    (void)e;
#pragma unroll
    for (int f = 0; f < F; ++f) gv[f][i] = far ? p.in[f][(size_t)nn * K + idx] : 0.0;
  }
#pragma unroll
  for (int i = 0; i < LPW; ++i) {
    const int j = wt + WL * i;
    if (j < Np) {
#pragma unroll
      for (int f = 0; f < F; ++f) us[tw][f][j][lane] = tl[f][i];
    }
  }
  __syncthreads();

  // ---- 2. numerical fluxes at the face nodes, scaled by face_J / J ---------
#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wt + WL * i;
    if (m >= NF) break;
    const int e = m / Nfp, d = m - e * Nfp;
    const int own = T::fnode(e, d);
    double a[F], b[F];
#pragma unroll
    for (int f = 0; f < F; ++f) a[f] = us[tw][f][own][lane];
    // outward normal and face_J/J of face e from the metric terms
    const double vx = e == 0 ? sx : (e == 1 ? rx + sx : rx);
    const double vy = e == 0 ? sy : (e == 1 ? ry + sy : ry);
    const double fs = sqrt(vx * vx + vy * vy);
    const double sg = (e == 1 ? 1.0 : -1.0) / (valid ? fs : 1.0);
    const double nx = sg * vx, ny = sg * vy;
    const int nf = code[i] & 3, idx = code[i] >> 2;
    if (nf == kBoundaryCode) {
      if (PHYS == kMaxwellTM) {
        // PEC wall: (Hx, Hy, Ez)+ = (Hx, Hy, -Ez)-
        b[0] = a[0];
        if (F == 3) {
          b[1 % F] = a[1 % F];
          b[2 % F] = -a[2 % F];
        }
      } else if (p.bc == kBcSinSin && valid) {
        const double xb = p.bx[idx + d], yb = p.by[idx + d];
        b[0] = sin((xb + 1.0 - t) * M_PI) * sin((yb + 1.0 - t) * M_PI);
      } else {
        b[0] = 0.0;
      }
    } else {
      const int nn = T::fnode(nf, Nfp - 1 - d);
      const int l = idx - c0;
      if ((unsigned)l < (unsigned)kTile) {
#pragma unroll
        for (int f = 0; f < F; ++f) b[f] = us[tw][f][nn][l];
      } else {
#pragma unroll
        for (int f = 0; f < F; ++f) b[f] = gv[f][i];
      }
    }
    if (PHYS == kMaxwellTM) {
      // Hx, Hy fluxes: (ny, -nx) * w / 2 with the face's normal
      const double dHx = a[0] - b[0];
      const double dHy = a[1 % F] - b[1 % F];
      const double sEz = a[2 % F] + b[2 % F];
      const double w = sEz + ny * dHx - nx * dHy;
      const double fez = 0.5 * (ny * (a[0] + b[0]) - nx * (a[1 % F] + b[1 % F]) + a[2 % F] - b[2 % F]) * fs;
      if (kFactor) {
        // the lift applies the normal per face: the tile keeps w / 2 once
        fl[tw][0][m][lane] = 0.5 * w * fs;
        fl[tw][1 % FL][m][lane] = fez;
      } else {
        fl[tw][0][m][lane] = 0.5 * ny * w * fs;
        fl[tw][1 % FL][m][lane] = -0.5 * nx * w * fs;
        fl[tw][2 % FL][m][lane] = fez;
      }
    } else {
      const double an = p.ax * nx + p.ay * ny;
      const double U = an >= 0.0 ? a[0] : b[0];
      fl[tw][0][m][lane] = U * an * fs;
    }
  }
  __syncthreads();

  // ---- 3. stage operands of this wave's rows (in flight during the FMAs, or
  //         loaded after the products where the code shape says so) ---------
  const int j0 = (part * WL + wt) * RPW;
  if (j0 >= Np) return;  // a wave without output rows (no barrier follows)
  double o1[F][RPW], o2[F][RPW], o3[F][RPW];
  auto load_operands = [&]() {
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    const int j = j0 + i;
    const bool ok = valid && j < Np;
    const size_t idx = (size_t)j * K + c;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      o1[f][i] = o2[f][i] = o3[f][i] = 0.0;
      if (ok) {
        if (STAGE == 2) {
          o1[f][i] = ld_once<NT>(p.acc[f] + idx);
          o2[f][i] = ld_once<NT>(p.u[f] + idx);
        } else if (STAGE == 3) {
          o2[f][i] = ld_once<NT>(p.u[f] + idx);
        } else if (STAGE == 4) {
          o1[f][i] = ld_once<NT>(p.acc[f] + idx);
          o2[f][i] = ld_once<NT>(p.u[f] + idx);
          o3[f][i] = ld_once<NT>(p.k3[f] + idx);
        }
      }
    }
  }

  };
  if constexpr (!kDelayOperands) load_operands();

  // ---- 4. volume term: Dr.u and Ds.u for this wave's rows -------------------
  double rr[F][RPW], ss[F][RPW];
#pragma unroll
  for (int f = 0; f < F; ++f)
#pragma unroll
    for (int i = 0; i < RPW; ++i) rr[f][i] = ss[f][i] = 0.0;
#pragma unroll kUnroll
  for (int k = 0; k < Np; ++k) {
    double uk[F];
#pragma unroll
    for (int f = 0; f < F; ++f) uk[f] = us[tw][f][k][lane];
    const double* dr = p.DrT + k * S + j0;
    const double* ds = p.DsT + k * S + j0;
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

  // ---- 5. lift: res -= LIFT . (flux * face_J / J) ----------------------------
  if constexpr (kFactor) {
    // per face e: sw = LIFT_e . (w / 2), then Hx -= ny_e sw, Hy += nx_e sw
    // (the face loop stays rolled where the code shape limits the unroll)
    constexpr int kFaceUnroll = kUnroll < 64 ? 1 : 3;
#pragma unroll kFaceUnroll
    for (int e = 0; e < 3; ++e) {
      double sw[RPW];
#pragma unroll
      for (int i = 0; i < RPW; ++i) sw[i] = 0.0;
#pragma unroll kUnroll
      for (int d = 0; d < Nfp; ++d) {
        const int m = e * Nfp + d;
        const double fw = fl[tw][0][m][lane];
        const double fe = fl[tw][1][m][lane];
        const double* lt = p.LT + m * S + j0;
#pragma unroll
        for (int i = 0; i < RPW; ++i) {
          const double cl = lt[i];
          sw[i] = fma(cl, fw, sw[i]);
          res[2 % F][i] = fma(-cl, fe, res[2 % F][i]);
        }
      }
      const double vx = e == 0 ? sx : (e == 1 ? rx + sx : rx);
      const double vy = e == 0 ? sy : (e == 1 ? ry + sy : ry);
      const double sg = (e == 1 ? 1.0 : -1.0) / (valid ? sqrt(vx * vx + vy * vy) : 1.0);
      const double nx = sg * vx, ny = sg * vy;
#pragma unroll
      for (int i = 0; i < RPW; ++i) {
        res[0][i] = fma(-ny, sw[i], res[0][i]);
        res[1 % F][i] = fma(nx, sw[i], res[1 % F][i]);
      }
    }
  } else {
#pragma unroll kUnroll
    for (int m = 0; m < NF; ++m) {
      double fm[F];
#pragma unroll
      for (int f = 0; f < F; ++f) fm[f] = fl[tw][f][m][lane];
      const double* lt = p.LT + m * S + j0;
#pragma unroll
      for (int i = 0; i < RPW; ++i) {
        const double cl = lt[i];
#pragma unroll
        for (int f = 0; f < F; ++f) res[f][i] = fma(-cl, fm[f], res[f][i]);
      }
    }
  }

  if constexpr (kDelayOperands) load_operands();
  // ---- 6. stage update (reference operand order: out = a * x + y) ------------
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
        st_once<NT>(p.acc[f] + idx, r);
        p.out[f][idx] = p.c1 * r + us[tw][f][j][lane];
      } else if (STAGE == 2) {
        st_once<NT>(p.acc[f] + idx, 2.0 * r + o1[f][i]);
        p.out[f][idx] = p.c1 * r + o2[f][i];
      } else if (STAGE == 3) {
        st_once<NT>(p.k3[f] + idx, r);
        p.out[f][idx] = p.c1 * r + o2[f][i];
      } else {
        const double w1 = 2.0 * o3[f][i] + r;
        const double w2 = p.c1 * o1[f][i] + o2[f][i];
        st_once<NT>(p.u[f] + idx, p.c1 * w1 + w2);
      }
    }
  }
}

}  // namespace dgc
