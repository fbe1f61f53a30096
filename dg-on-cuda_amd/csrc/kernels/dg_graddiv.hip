// Weak-form gradient and divergence with central fluxes on the GPU, and their
// C ABI (include/dgc_hip.h: dgc_graddiv_*, dgc_grad2d_apply, dgc_div2d_apply).
//
//   grad: (gx, gy) = -(Dx u, Dy u) + LIFT (face_J/J) {u} (nx, ny)
//         src/grad_2d.h:95-187
//   div:  d = -(Dx u0 + Dy u1) + LIFT (face_J/J) ({u0} nx + {u1} ny)
//         src/div_2d.h:95-188
// with {u} = (u- + u+)/2 and, on boundary faces, the exterior value
// u+ = g[b] + beta u- at boundary node b (numbered cell by cell, face by face,
// as the reference calls its BC object), which covers the affine boundary
// conditions of the Stokes splitting (examples/stokes_2d/stokes_2d.h:100-183).
//
// WT waves per tile of 64 consecutive cells (lane = cell), 4 / WT tiles per
// 256-thread workgroup, as the RK stage kernels: the tile's node rows are
// staged in LDS, the tile's waves split the face nodes for the fluxes and the
// node rows for the volume and lift products. Fewer waves per tile give each
// wave more rows in flight. Both operators are HBM-bound: per cell they read the input
// rows, four metric terms and three connectivity words and write the output
// rows once.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../../include/dgc_hip.h"
#include "dg_internal.hpp"

#define DGC_HIP_TRY(call)                                                                     \
  do {                                                                                        \
    hipError_t err_ = (call);                                                                 \
    if (err_ != hipSuccess)                                                                   \
      return dgc_set_error(DGC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(err_)); \
  } while (0)

#define DGC_REQUIRE(cond, msg) \
  do {                         \
    if (!(cond)) return dgc_set_error(DGC_ERR_INVALID, msg); \
  } while (0)

namespace dgc {
namespace {

struct GradDivParams {
  int K;
  const double* __restrict__ DrT;  // [Np][S] transposed, rows padded to S
  const double* __restrict__ DsT;
  const double* __restrict__ LT;   // [3 Nfp][S]
  const double* __restrict__ geo;
  const int* __restrict__ nbr;
  const double* in0;
  const double* in1;  // div only
  const double* g0;   // boundary data (null: 0)
  const double* g1;   // div only
  double beta;
  double* out0;  // grad: x component; div: the divergence
  double* out1;  // grad: y component
};

// LDS bytes of a workgroup: node rows and face fluxes of its 4 / WT tiles
constexpr int graddiv_lds_bytes(int N, bool div, int wt) {
  return (kWaves / wt) * ((div ? 2 : 1) * (N + 1) * (N + 2) / 2 + (div ? 1 : 2) * 3 * (N + 1)) * kTile * 8;
}
// a tiling is built when its LDS fits two workgroups per CU (WT = 4 always)
constexpr bool graddiv_fits(int N, bool div, int wt) { return wt == 4 || graddiv_lds_bytes(N, div, wt) <= 64 * 1024; }

// DIV = false: gradient of in0 into (out0, out1); true: divergence of (in0, in1) into out0.
template <int N, bool DIV, int WT>
__global__ __launch_bounds__(kBlock) void graddiv_kernel(const GradDivParams p) {
  using T = Tri<N>;
  constexpr int Np = T::Np, Nfp = T::Nfp, NF = 3 * Nfp;
  constexpr int FI = DIV ? 2 : 1;  // input fields
  constexpr int FO = DIV ? 1 : 2;  // flux / output fields
  constexpr int TPW = kWaves / WT;  // tiles per workgroup
  constexpr int RPW = (Np + WT - 1) / WT;
  constexpr int FPW = (NF + WT - 1) / WT;
  constexpr int S = kWaves * ((Np + kWaves - 1) / kWaves);  // operator row stride
  __shared__ double lds_us[TPW][FI][Np][kTile];
  __shared__ double lds_fl[TPW][FO][NF][kTile];

  const int K = p.K;
  const int lane = threadIdx.x & (kTile - 1);
  const int wave = uniform(threadIdx.x >> 6);
  const int tw = wave / WT, wt = wave - tw * WT;
  auto& us = lds_us[tw];
  auto& fl = lds_fl[tw];
  const int c0 = (xcd_group(blockIdx.x, gridDim.x) * TPW + tw) * kTile;
  const int c = c0 + lane;
  const bool valid = c < K;
  const double* in[2] = {p.in0, p.in1};
  const double* g[2] = {p.g0, p.g1};

  // ---- 1. loads: node rows (strided over the waves), metric, neighbour codes
  int code[FPW];
#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wt + WT * i;
    code[i] = (valid && m < NF) ? p.nbr[(size_t)(m / Nfp) * K + c] : kBoundaryCode;
  }
  double rx = 0.0, ry = 0.0, sx = 0.0, sy = 0.0;
  if (valid) {
    rx = p.geo[(size_t)kRx * K + c];
    ry = p.geo[(size_t)kRy * K + c];
    sx = p.geo[(size_t)kSx * K + c];
    sy = p.geo[(size_t)kSy * K + c];
  }
  double tl[FI][RPW];
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    const int j = wt + WT * i;
#pragma unroll
    for (int f = 0; f < FI; ++f) tl[f][i] = (valid && j < Np) ? in[f][(size_t)j * K + c] : 0.0;
  }
  // neighbour traces outside the tile, in flight with the rows
  double gv[FI][FPW];
#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int d = (wt + WT * i) % Nfp;
    const int nf = code[i] & 3, idx = code[i] >> 2;
    const bool far = nf != kBoundaryCode && (unsigned)(idx - c0) >= (unsigned)kTile;
    const int nn = T::fnode(nf, Nfp - 1 - d);
#pragma unroll
    for (int f = 0; f < FI; ++f) gv[f][i] = far ? in[f][(size_t)nn * K + idx] : 0.0;
  }
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    const int j = wt + WT * i;
    if (j < Np) {
#pragma unroll
      for (int f = 0; f < FI; ++f) us[f][j][lane] = tl[f][i];
    }
  }
  __syncthreads();

  // ---- 2. central fluxes at the face nodes, scaled by face_J / J
#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wt + WT * i;
    if (m >= NF) break;
    const int e = m / Nfp, d = m - e * Nfp;
    const int own = T::fnode(e, d);
    double a[FI], b[FI];
#pragma unroll
    for (int f = 0; f < FI; ++f) a[f] = us[f][own][lane];
    // outward normal and face_J/J of face e from the metric terms
    const double vx = e == 0 ? sx : (e == 1 ? rx + sx : rx);
    const double vy = e == 0 ? sy : (e == 1 ? ry + sy : ry);
    const double fs = sqrt(vx * vx + vy * vy);
    const double sg = (e == 1 ? 1.0 : -1.0) / (valid ? fs : 1.0);
    const double nx = sg * vx, ny = sg * vy;
    const int nf = code[i] & 3, idx = code[i] >> 2;
    if (nf == kBoundaryCode) {
#pragma unroll
      for (int f = 0; f < FI; ++f) b[f] = (valid && g[f] != nullptr ? g[f][idx + d] : 0.0) + p.beta * a[f];
    } else {
      const int nn = T::fnode(nf, Nfp - 1 - d);
      const int l = idx - c0;
      if ((unsigned)l < (unsigned)kTile) {
#pragma unroll
        for (int f = 0; f < FI; ++f) b[f] = us[f][nn][l];
      } else {
#pragma unroll
        for (int f = 0; f < FI; ++f) b[f] = gv[f][i];
      }
    }
    if (DIV) {
      fl[0][m][lane] = 0.5 * ((a[0] + b[0]) * nx + (a[FI - 1] + b[FI - 1]) * ny) * fs;
    } else {
      const double fc = 0.5 * (a[0] + b[0]) * fs;
      fl[0][m][lane] = fc * nx;
      fl[FO - 1][m][lane] = fc * ny;
    }
  }
  __syncthreads();

  // ---- 3. volume term: Dr.u and Ds.u for this wave's rows
  const int j0 = wt * RPW;
  double rr[FI][RPW], ss[FI][RPW];
#pragma unroll
  for (int f = 0; f < FI; ++f)
#pragma unroll
    for (int i = 0; i < RPW; ++i) rr[f][i] = ss[f][i] = 0.0;
#pragma unroll
  for (int k = 0; k < Np; ++k) {
    double uk[FI];
#pragma unroll
    for (int f = 0; f < FI; ++f) uk[f] = us[f][k][lane];
    const double* dr = p.DrT + k * S + j0;
    const double* ds = p.DsT + k * S + j0;
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      const double cr = dr[i], cs = ds[i];
#pragma unroll
      for (int f = 0; f < FI; ++f) {
        rr[f][i] = fma(cr, uk[f], rr[f][i]);
        ss[f][i] = fma(cs, uk[f], ss[f][i]);
      }
    }
  }
  double res[FO][RPW];
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    if (DIV) {
      res[0][i] = -(rx * rr[0][i] + sx * ss[0][i]) - (ry * rr[FI - 1][i] + sy * ss[FI - 1][i]);
    } else {
      res[0][i] = -(rx * rr[0][i] + sx * ss[0][i]);
      res[FO - 1][i] = -(ry * rr[0][i] + sy * ss[0][i]);
    }
  }

  // ---- 4. lift of the scaled fluxes
#pragma unroll
  for (int m = 0; m < NF; ++m) {
    double fm[FO];
#pragma unroll
    for (int f = 0; f < FO; ++f) fm[f] = fl[f][m][lane];
    const double* lt = p.LT + m * S + j0;
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      const double cl = lt[i];
#pragma unroll
      for (int f = 0; f < FO; ++f) res[f][i] = fma(cl, fm[f], res[f][i]);
    }
  }

  if (!valid) return;
  double* out[2] = {p.out0, p.out1};
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    const int j = j0 + i;
    if (j >= Np) break;
#pragma unroll
    for (int f = 0; f < FO; ++f) out[f][(size_t)j * K + c] = res[f][i];
  }
}

using GradDivFn = void (*)(GradDivParams);

template <int N, bool DIV>
GradDivFn pick_wt(int wt) {
  if constexpr (graddiv_fits(N, DIV, 1)) {
    if (wt == 1) return graddiv_kernel<N, DIV, 1>;
  }
  if constexpr (graddiv_fits(N, DIV, 2)) {
    if (wt == 2) return graddiv_kernel<N, DIV, 2>;
  }
  return wt == 4 ? graddiv_kernel<N, DIV, 4> : nullptr;
}

template <bool DIV>
GradDivFn pick(int order, int wt) {
  switch (order) {
    case 1: return pick_wt<1, DIV>(wt);
    case 2: return pick_wt<2, DIV>(wt);
    case 3: return pick_wt<3, DIV>(wt);
    case 4: return pick_wt<4, DIV>(wt);
    case 5: return pick_wt<5, DIV>(wt);
    case 6: return pick_wt<6, DIV>(wt);
    case 7: return pick_wt<7, DIV>(wt);
    default: return nullptr;
  }
}

}  // namespace
}  // namespace dgc

struct dgc_graddiv_s {
  dgc_disc_s* disc = nullptr;
  double *DrT = nullptr, *DsT = nullptr, *LT = nullptr;
  int wt = 0;  // waves per tile; 0: the default for the order
};

extern "C" {

int dgc_graddiv_create(dgc_disc_t disc, const double* Dr, const double* Ds, const double* lift,
                       dgc_graddiv_t* out) {
  DGC_REQUIRE(disc != nullptr && out != nullptr, "dgc_graddiv_create: null argument");
  DGC_REQUIRE(Dr && Ds && lift, "dgc_graddiv_create: null operator");
  const int Np = disc->Np, NF = 3 * disc->Nfp, S = dgc::padded_rows(Np);
  std::vector<double> drT(size_t(Np) * S, 0.0), dsT(size_t(Np) * S, 0.0), lT(size_t(NF) * S, 0.0);
  for (int j = 0; j < Np; ++j) {
    for (int k = 0; k < Np; ++k) {
      drT[size_t(k) * S + j] = Dr[size_t(j) * Np + k];
      dsT[size_t(k) * S + j] = Ds[size_t(j) * Np + k];
    }
    for (int m = 0; m < NF; ++m) lT[size_t(m) * S + j] = lift[size_t(j) * NF + m];
  }
  DGC_HIP_TRY(hipSetDevice(disc->ctx->device));
  auto* op = new dgc_graddiv_s;
  op->disc = disc;
  auto fail = [&](hipError_t e) {
    (void)hipFree(op->DrT);
    (void)hipFree(op->DsT);
    (void)hipFree(op->LT);
    delete op;
    return dgc_set_error(DGC_ERR_HIP, std::string("dgc_graddiv_create: ") + hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = hipMalloc(&op->DrT, drT.size() * sizeof(double))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&op->DsT, dsT.size() * sizeof(double))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&op->LT, lT.size() * sizeof(double))) != hipSuccess) return fail(e);
  if ((e = hipMemcpy(op->DrT, drT.data(), drT.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e);
  if ((e = hipMemcpy(op->DsT, dsT.data(), dsT.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e);
  if ((e = hipMemcpy(op->LT, lT.data(), lT.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e);
  *out = op;
  return DGC_OK;
}

int dgc_graddiv_destroy(dgc_graddiv_t op) {
  if (op == nullptr) return DGC_OK;
  (void)hipSetDevice(op->disc->ctx->device);
  (void)hipStreamSynchronize(op->disc->ctx->stream);
  (void)hipFree(op->DrT);
  (void)hipFree(op->DsT);
  (void)hipFree(op->LT);
  delete op;
  return DGC_OK;
}

static int graddiv_launch(dgc_graddiv_t op, bool div, const double* u0, const double* u1, const double* g0,
                          const double* g1, double beta, double* out0, double* out1, const char* name) {
  const dgc_disc_s* d = op->disc;
  dgc::GradDivParams p{};
  p.K = d->K;
  p.DrT = op->DrT;
  p.DsT = op->DsT;
  p.LT = op->LT;
  p.geo = d->geo;
  p.nbr = d->nbr;
  p.in0 = u0;
  p.in1 = u1;
  p.g0 = g0;
  p.g1 = g1;
  p.beta = beta;
  p.out0 = out0;
  p.out1 = out1;
  // default tiling per order, measured at K = 2.1M (1 wave per tile at N <= 2:
  // 6.2 / 3.8 TB/s against 3.6 / 2.9 with 4; 2 at N = 3 and 5; N >= 6 only
  // fits 4 within 64 KB of LDS)
  static constexpr int kDefaultWt[8] = {4, 1, 1, 2, 4, 2, 4, 4};
  int wt = op->wt != 0 ? op->wt : kDefaultWt[d->order];
  dgc::GradDivFn fn = div ? dgc::pick<true>(d->order, wt) : dgc::pick<false>(d->order, wt);
  if (fn == nullptr) {  // the tiling is not built for this order: one tile per workgroup
    wt = 4;
    fn = div ? dgc::pick<true>(d->order, wt) : dgc::pick<false>(d->order, wt);
  }
  if (fn == nullptr) return dgc_set_error(DGC_ERR_UNSUPPORTED, std::string(name) + ": order not built");
  DGC_HIP_TRY(hipSetDevice(d->ctx->device));
  const int tiles = (d->K + dgc::kTile - 1) / dgc::kTile;
  const int tpw = dgc::kWaves / wt;
  hipLaunchKernelGGL(fn, dim3((tiles + tpw - 1) / tpw), dim3(dgc::kBlock), 0, d->ctx->stream, p);
  DGC_HIP_TRY(hipGetLastError());
  return DGC_OK;
}

int dgc_graddiv_set_tiling(dgc_graddiv_t op, int waves_per_tile) {
  DGC_REQUIRE(op != nullptr, "dgc_graddiv_set_tiling: null operator");
  DGC_REQUIRE(waves_per_tile == 0 || waves_per_tile == 1 || waves_per_tile == 2 || waves_per_tile == 4,
              "dgc_graddiv_set_tiling: waves per tile must be 0 (default), 1, 2 or 4");
  op->wt = waves_per_tile;
  return DGC_OK;
}

int dgc_grad2d_apply(dgc_graddiv_t op, const double* u, const double* g, double beta, double* out_x,
                     double* out_y) {
  DGC_REQUIRE(op != nullptr && u && out_x && out_y, "dgc_grad2d_apply: null argument");
  DGC_REQUIRE(out_x != u && out_y != u && out_x != out_y, "dgc_grad2d_apply: outputs must not alias the input");
  return graddiv_launch(op, false, u, nullptr, g, nullptr, beta, out_x, out_y, "dgc_grad2d_apply");
}

int dgc_div2d_apply(dgc_graddiv_t op, const double* u0, const double* u1, const double* g0, const double* g1,
                    double beta, double* out) {
  DGC_REQUIRE(op != nullptr && u0 && u1 && out, "dgc_div2d_apply: null argument");
  DGC_REQUIRE(out != u0 && out != u1, "dgc_div2d_apply: the output must not alias an input");
  return graddiv_launch(op, true, u0, u1, g0, g1, beta, out, nullptr, "dgc_div2d_apply");
}

}  // extern "C"
