// 1D linear advection of examples/advection_1d on the GPU, with the fused
// RK4 stage update of the 2D path (include/dgc_hip.h: dgc_adv1d_*).
//
//   u_t + a u_x = 0 on [0, 2 pi], a = 2 pi, K uniform cells of order N,
//   inflow u(0, t) = -sin(2 pi t), outflow copies the interior value,
//   upwind flux F(l, r) = a (l + r) / 2 + a (l - r) / 2
//   (examples/advection_1d/advection_1d.h:38-199, d_advection_1d.cuh:37-98)
// strong form per cell: out = -a D u + L v with v = (F_left - a u_0, 0, ...,
// a u_N - F_right), D = Dr (2/h), L = V V^T (2/h).
//
// Node-major layout u[j * K + c] as the reference's GPU path; one thread per
// cell (coalesced across the cells of a wave). The stages follow
// src/explicit_runge_kutta.h:33-51 with the operand order of its axpys, as
// the 2D stage kernels do.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../../include/dgc_hip.h"
#include "../host/refelem.hpp"
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

namespace {

constexpr int kMaxOrder1D = 12;
constexpr double kWave = 2.0 * M_PI;

struct Adv1dParams {
  int K;
  const double* __restrict__ D;  // [(N+1)^2] row-major
  const double* __restrict__ L;
  double t;     // stage time
  double c1;    // stage coefficient
  const double* in;   // stage input
  double* out;        // stage 0: the right-hand side; 1-3: the next stage input
  double* u;          // the step's state (stage 4 writes it)
  double* acc;        // k1, then 2 k2 + k1
  double* k3;
};

__device__ inline double upwind(double l, double r) {
  return kWave * 0.5 * (l + r) + kWave * 0.5 * (1.0 - 0.0) * (l - r);
}

// STAGE 0: out = rhs(in); 1..4: the RK4 stages of src/explicit_runge_kutta.h:33-51
//   1: acc = k1, out = c1 k1 + u        2: acc = 2 k2 + k1, out = c1 k2 + u
//   3: k3 saved, out = c1 k3 + u        4: u = c1 (2 k3 + k4) + (c1 acc + u)
template <int N, int STAGE>
__global__ __launch_bounds__(256) void adv1d_kernel(const Adv1dParams p) {
  constexpr int Np = N + 1;
  const int K = p.K;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= K) return;
  double v[Np];
#pragma unroll
  for (int j = 0; j < Np; ++j) v[j] = p.in[(size_t)j * K + c];
  const double aL = c == 0 ? -sin(2.0 * M_PI * p.t) : p.in[(size_t)N * K + c - 1];
  const double bR = c == K - 1 ? v[N] : p.in[c + 1];
  const double f0 = upwind(aL, v[0]) - kWave * v[0];
  const double f1 = kWave * v[N] - upwind(v[N], bR);
#pragma unroll
  for (int j = 0; j < Np; ++j) {
    double dv = 0.0;
#pragma unroll
    for (int k = 0; k < Np; ++k) dv += p.D[j * Np + k] * v[k];
    const double r = -kWave * dv + (p.L[j * Np] * f0 + p.L[j * Np + N] * f1);
    const size_t i = (size_t)j * K + c;
    if (STAGE == 0) {
      p.out[i] = r;
    } else if (STAGE == 1) {
      p.acc[i] = r;
      p.out[i] = p.c1 * r + p.u[i];
    } else if (STAGE == 2) {
      p.acc[i] = 2.0 * r + p.acc[i];
      p.out[i] = p.c1 * r + p.u[i];
    } else if (STAGE == 3) {
      p.k3[i] = r;
      p.out[i] = p.c1 * r + p.u[i];
    } else {
      const double w1 = 2.0 * p.k3[i] + r;
      const double w2 = p.c1 * p.acc[i] + p.u[i];
      p.u[i] = p.c1 * w1 + w2;
    }
  }
}

using Adv1dFn = void (*)(Adv1dParams);

template <int N>
Adv1dFn pick_stage(int stage) {
  switch (stage) {
    case 0: return adv1d_kernel<N, 0>;
    case 1: return adv1d_kernel<N, 1>;
    case 2: return adv1d_kernel<N, 2>;
    case 3: return adv1d_kernel<N, 3>;
    default: return adv1d_kernel<N, 4>;
  }
}

Adv1dFn pick(int order, int stage) {
  switch (order) {
    case 1: return pick_stage<1>(stage);
    case 2: return pick_stage<2>(stage);
    case 3: return pick_stage<3>(stage);
    case 4: return pick_stage<4>(stage);
    case 5: return pick_stage<5>(stage);
    case 6: return pick_stage<6>(stage);
    case 7: return pick_stage<7>(stage);
    case 8: return pick_stage<8>(stage);
    case 9: return pick_stage<9>(stage);
    case 10: return pick_stage<10>(stage);
    case 11: return pick_stage<11>(stage);
    case 12: return pick_stage<12>(stage);
    default: return nullptr;
  }
}

}  // namespace

struct dgc_adv1d_s {
  dgc_ctx_s* ctx = nullptr;
  int K = 0, order = 0;
  double *D = nullptr, *L = nullptr;
  double* work = nullptr;  // acc, k3 and two stage inputs: [4][K (order+1)]
};

extern "C" {

int dgc_adv1d_operators(int num_cells, int order, double* D, double* L, double* x) {
  DGC_REQUIRE(num_cells > 0 && order >= 1 && order <= kMaxOrder1D, "dgc_adv1d_operators: bad size or order");
  const int Np = order + 1;
  std::vector<double> r, w;
  dgc_host::lgl_points(Np, r, w);
  dgc_host::Mat V(Np, Np), Vr(Np, Np);
  for (int i = 0; i < Np; ++i)
    for (int j = 0; j < Np; ++j) {
      V(i, j) = dgc_host::jacobi_p(r[i], 0.0, 0.0, j);
      Vr(i, j) = dgc_host::grad_jacobi_p(r[i], 0.0, 0.0, j);
    }
  const double h = 2.0 * M_PI / num_cells;
  const dgc_host::Mat Dr = dgc_host::matmul(Vr, dgc_host::inverse(V));
  const dgc_host::Mat Mi = dgc_host::matmul(V, dgc_host::transpose(V));
  for (int i = 0; i < Np * Np; ++i) {
    if (D) D[i] = Dr.a[i] * (2.0 / h);
    if (L) L[i] = Mi.a[i] * (2.0 / h);
  }
  // node positions, node-major (x[j * K + c]) as advection_1d::initialize_dofs
  if (x)
    for (int j = 0; j < Np; ++j)
      for (int c = 0; c < num_cells; ++c) x[size_t(j) * num_cells + c] = c * h + (1.0 + r[j]) * 0.5 * h;
  return DGC_OK;
}

int dgc_adv1d_create(dgc_ctx_t ctx, int num_cells, int order, const double* D, const double* L, dgc_adv1d_t* out) {
  DGC_REQUIRE(ctx != nullptr && out != nullptr, "dgc_adv1d_create: null argument");
  DGC_REQUIRE(num_cells > 0 && order >= 1, "dgc_adv1d_create: bad size or order");
  if (order > kMaxOrder1D)
    return dgc_set_error(DGC_ERR_UNSUPPORTED,
                         "dgc_adv1d_create: order must be in [1, " + std::to_string(kMaxOrder1D) + "]");
  const int Np = order + 1;
  std::vector<double> hD(Np * Np), hL(Np * Np);
  if (D && L) {
    hD.assign(D, D + Np * Np);
    hL.assign(L, L + Np * Np);
  } else {
    int rc = dgc_adv1d_operators(num_cells, order, hD.data(), hL.data(), nullptr);
    if (rc) return rc;
  }
  DGC_HIP_TRY(hipSetDevice(ctx->device));
  auto* op = new dgc_adv1d_s;
  op->ctx = ctx;
  op->K = num_cells;
  op->order = order;
  const size_t n = size_t(num_cells) * Np;
  if (hipMalloc(&op->D, hD.size() * sizeof(double)) != hipSuccess ||
      hipMalloc(&op->L, hL.size() * sizeof(double)) != hipSuccess ||
      hipMalloc(&op->work, 4 * n * sizeof(double)) != hipSuccess ||
      hipMemcpy(op->D, hD.data(), hD.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(op->L, hL.data(), hL.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    dgc_adv1d_destroy(op);
    return dgc_set_error(DGC_ERR_HIP, "dgc_adv1d_create: device allocation or copy failed");
  }
  *out = op;
  return DGC_OK;
}

int dgc_adv1d_destroy(dgc_adv1d_t op) {
  if (op == nullptr) return DGC_OK;
  (void)hipStreamSynchronize(op->ctx->stream);
  (void)hipFree(op->D);
  (void)hipFree(op->L);
  (void)hipFree(op->work);
  delete op;
  return DGC_OK;
}

static int adv1d_launch(dgc_adv1d_t op, int stage, Adv1dParams& p) {
  Adv1dFn fn = pick(op->order, stage);
  if (fn == nullptr) return dgc_set_error(DGC_ERR_UNSUPPORTED, "dgc_adv1d: order not built");
  hipLaunchKernelGGL(fn, dim3((op->K + 255) / 256), dim3(256), 0, op->ctx->stream, p);
  DGC_HIP_TRY(hipGetLastError());
  return DGC_OK;
}

int dgc_adv1d_rhs(dgc_adv1d_t op, const double* in, double t, double* out) {
  DGC_REQUIRE(op && in && out && in != out, "dgc_adv1d_rhs: bad argument");
  Adv1dParams p{};
  p.K = op->K;
  p.D = op->D;
  p.L = op->L;
  p.t = t;
  p.in = in;
  p.out = out;
  return adv1d_launch(op, 0, p);
}

int dgc_adv1d_rk4_steps(dgc_adv1d_t op, double* u, double t0, double dt, int nsteps) {
  DGC_REQUIRE(op && u && nsteps >= 0, "dgc_adv1d_rk4_steps: bad argument");
  const size_t n = size_t(op->K) * (op->order + 1);
  Adv1dParams p{};
  p.K = op->K;
  p.D = op->D;
  p.L = op->L;
  p.u = u;
  p.acc = op->work;
  p.k3 = op->work + n;
  // the stage inputs alternate between two buffers: a stage reads its
  // neighbours' edge values while other threads write their own output
  double* w[2] = {op->work + 2 * n, op->work + 3 * n};
  double t = t0;
  for (int s = 0; s < nsteps; ++s) {
    const double ts[4] = {t, t + 0.5 * dt, t + 0.5 * dt, t + dt};
    const double cs[4] = {0.5 * dt, 0.5 * dt, dt, dt / 6.0};
    for (int stage = 1; stage <= 4; ++stage) {
      p.t = ts[stage - 1];
      p.c1 = cs[stage - 1];
      p.in = stage == 1 ? u : w[stage % 2];
      p.out = w[(stage + 1) % 2];
      int rc = adv1d_launch(op, stage, p);
      if (rc) return rc;
    }
    t += dt;
  }
  return DGC_OK;
}

}  // extern "C"
