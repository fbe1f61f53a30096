// C ABI of the IPDG operator and its solvers (include/dgc_hip.h):
// apply (the laplace_2d replacement), block-Jacobi PCG in FP64 and the
// FP32-inner / FP64-outer iterative-refinement variant.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../../include/dgc_hip.h"
#include "../host/refelem.hpp"
#include "dg_internal.hpp"
#include "dg_ipdg.hpp"
#include "dg_mpnn.hpp"
#include "dgc_ipdg_internal.hpp"

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
#define DGC_LAUNCH(call)                                                                        \
  do {                                                                                          \
    hipError_t err_ = (call);                                                                   \
    if (err_ != hipSuccess)                                                                     \
      return dgc_set_error(DGC_ERR_LAUNCH, std::string("failed to launch kernel of device function phase! ") + \
                                               hipGetErrorString(err_));                        \
  } while (0)

namespace {

template <typename T>
int alloc(T** p, size_t n) {
  DGC_HIP_TRY(hipMalloc(p, std::max<size_t>(n, 1) * sizeof(T)));
  return DGC_OK;
}

template <typename T>
int upload(T** p, const std::vector<T>& h) {
  int rc = alloc(p, h.size());
  if (rc) return rc;
  DGC_HIP_TRY(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return DGC_OK;
}

// Transposed, row-padded operator layout read by the tile kernels.
std::vector<double> transpose_pad(const double* A, int rows, int cols) {
  const int Npad = dgc::ipdg_rows(rows);
  std::vector<double> t(size_t(cols) * Npad, 0.0);
  for (int j = 0; j < rows; ++j)
    for (int k = 0; k < cols; ++k) t[size_t(k) * Npad + j] = A[size_t(j) * cols + k];
  return t;
}

template <typename T>
std::vector<T> cast(const std::vector<double>& v) {
  return std::vector<T>(v.begin(), v.end());
}

__global__ void axpby_kernel(double a, const double* x, double b, const double* y, double* out, size_t n) {
  const size_t stride = size_t(blockDim.x) * gridDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = a * x[i] + b * y[i];
}

// out32 = float(s * x64)
__global__ void to_float_kernel(const double* x, double s, float* out, size_t n) {
  const size_t stride = size_t(blockDim.x) * gridDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = float(s * x[i]);
}

// x64 += s * double(e32)
__global__ void add_scaled_float_kernel(double* x, double s, const float* e, size_t n) {
  const size_t stride = size_t(blockDim.x) * gridDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) x[i] += s * double(e[i]);
}

// probe vector: 1 at node k of every cell of colour `col`, 0 elsewhere
__global__ void probe_fill_kernel(const int* color, int col, int k, int Np, int K, double* u) {
  const size_t n = size_t(Np) * K;
  const size_t stride = size_t(blockDim.x) * gridDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int j = int(i / K), c = int(i % K);
    u[i] = (j == k && color[c] == col) ? 1.0 : 0.0;
  }
}

// blocks[(j*Np + k) * K + c] = out[j*K + c] for the cells of colour `col`
__global__ void probe_extract_kernel(const int* color, int col, int k, int Np, int K, const double* out,
                                     double* blocks) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= K || color[c] != col) return;
  for (int j = 0; j < Np; ++j) blocks[(size_t(j) * Np + k) * K + c] = out[size_t(j) * K + c];
}

unsigned grid_for(size_t n) { return unsigned(std::min<size_t>((n + 255) / 256, 8192)); }

}  // namespace

// ---------------------------------------------------------------------------
template <typename T>
int IpdgLevel<T>::init(const dgc_ipdg_s& s, const std::vector<std::vector<double>>& tables,
                       const std::vector<double>& geo, const std::vector<double>& J,
                       const std::vector<double>& tau) {
  const size_t n = size_t(s.disc->K) * s.disc->Np;
  const size_t nf = size_t(s.disc->K) * 3 * s.disc->Nfp;
  int rc;
  T** dst[7] = {&DrT, &DsT, &LT, &MT, &MDrT, &MDsT, &MLT};
  for (int i = 0; i < 7; ++i)
    if ((rc = upload(dst[i], cast<T>(tables[i])))) return rc;
  if ((rc = upload(&tau_d, cast<T>(tau)))) return rc;
  if (std::is_same<T, double>::value) {
    geo_d = reinterpret_cast<T*>(s.disc->geo);
    J_d = reinterpret_cast<T*>(s.disc->J);
    own_geo = false;
  } else {
    if ((rc = upload(&geo_d, cast<T>(geo))) || (rc = upload(&J_d, cast<T>(J)))) return rc;
    own_geo = true;
  }
  if ((rc = alloc(&qx, n)) || (rc = alloc(&qy, n)) || (rc = alloc(&du, nf)) || (rc = alloc(&r, n)) ||
      (rc = alloc(&z, n)) || (rc = alloc(&p, n)) || (rc = alloc(&p2, n)) || (rc = alloc(&q, n)) ||
      (rc = alloc(&x, n)) || (rc = alloc(&partial, 2 * size_t(s.grid))) || (rc = alloc(&dpartial, size_t(s.grid))) ||
      (rc = alloc(&scal, dgc::kCgScalars)) || (rc = alloc(&rec, 2 * dgc::kRecStride)))
    return rc;
  return DGC_OK;
}

template <typename T>
void IpdgLevel<T>::release() {
  (void)hipFree(DrT);
  (void)hipFree(DsT);
  (void)hipFree(LT);
  (void)hipFree(MT);
  (void)hipFree(MDrT);
  (void)hipFree(MDsT);
  (void)hipFree(MLT);
  (void)hipFree(tau_d);
  if (own_geo) {
    (void)hipFree(geo_d);
    (void)hipFree(J_d);
  }
  (void)hipFree(qx);
  (void)hipFree(qy);
  (void)hipFree(du);
  (void)hipFree(r);
  (void)hipFree(z);
  (void)hipFree(p);
  (void)hipFree(p2);
  (void)hipFree(q);
  (void)hipFree(x);
  (void)hipFree(partial);
  (void)hipFree(dpartial);
  (void)hipFree(scal);
  (void)hipFree(rec);
  (void)hipFree(P);
  if (graph) (void)hipGraphExecDestroy(graph);
  *this = IpdgLevel<T>();
}

template <typename T>
dgc::IpdgParams<T> IpdgLevel<T>::params(const dgc_ipdg_s& s) const {
  dgc::IpdgParams<T> pr{};
  pr.K = s.disc->K;
  pr.DrT = DrT;
  pr.DsT = DsT;
  pr.LT = LT;
  pr.MT = MT;
  pr.MDrT = MDrT;
  pr.MDsT = MDsT;
  pr.MLT = MLT;
  pr.geo = geo_d;
  pr.J = J_d;
  pr.tau = tau_d;
  pr.nbr = s.disc->nbr;
  pr.neumann = s.neumann;
  pr.shift = T(s.shift);
  pr.qx = qx;
  pr.qy = qy;
  pr.du = du;
  return pr;
}

template struct IpdgLevel<double>;
template struct IpdgLevel<float>;

// The fused iteration (3 kernels per block-Jacobi iteration: the scalar
// reductions are finished inside the kernels that consume them) runs when
// the partial-sum count is small enough for every workgroup to sum them all.
static bool fused_iteration(const dgc_ipdg_s* s) { return s->grid <= dgc::kFusedMaxParts; }

// Enqueue CG initialisation: r = b - S x (x given), z = P r and the partials
// (r.z, r.r); p = 0 with beta = 0, so the first iteration's direction is z.
// Unfused: the scalar kernel finishes the initial scalars. Fused: record 0
// marks a fresh initialisation (kPhase) for the first iteration's pass 1, and
// record 1 carries the done flag that makes the update form z only.
template <typename T>
static int cg_init(dgc_ipdg_s* s, IpdgLevel<T>& L, const T* b, double tol_rel2_times_bb, int maxit,
                   bool x_is_zero, double* history) {
  hipStream_t st = s->disc->ctx->stream;
  const size_t n = size_t(s->disc->K) * s->disc->Np;
  const bool fused = fused_iteration(s);
  T* init_scal = L.scal;
  if (fused) {
    T h[2 * dgc::kRecStride] = {};
    h[dgc::kTol2] = T(tol_rel2_times_bb);
    h[dgc::kMaxit] = T(maxit);
    h[dgc::kPhase] = T(1);
    h[dgc::kRecStride + dgc::kDone] = T(1);
    DGC_HIP_TRY(hipMemcpyAsync(L.rec, h, sizeof(h), hipMemcpyHostToDevice, st));
    init_scal = L.rec + dgc::kRecStride;
  } else {
    // done = 1 while initialising: cg_update then leaves x and r alone and
    // only forms z = P r and the partial sums; the kInit reduction clears it
    T h[dgc::kCgScalars] = {0, 0, 0, 0, 0, 1, T(tol_rel2_times_bb), T(maxit)};
    DGC_HIP_TRY(hipMemcpyAsync(L.scal, h, sizeof(h), hipMemcpyHostToDevice, st));
  }
  if (x_is_zero) {
    DGC_HIP_TRY(hipMemcpyAsync(L.r, b, n * sizeof(T), hipMemcpyDeviceToDevice, st));
  } else {
    dgc::IpdgParams<T> pr = L.params(*s);
    pr.u = L.x;
    pr.out = L.q;
    DGC_LAUNCH(dgc::ipdg_apply<T>(s->disc->order, pr, dgc::kSpd, st));
    // r = b - q  (kept in T; for T = double this is the FP64 residual)
    if (std::is_same<T, double>::value) {
      hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(256), 0, st, 1.0, (const double*)b, -1.0,
                         (const double*)L.q, (double*)L.r, n);
      DGC_LAUNCH(hipGetLastError());
    } else {
      return dgc_set_error(DGC_ERR_INVALID, "cg_init: non-zero initial guess needs FP64");
    }
  }
  DGC_HIP_TRY(hipMemsetAsync(L.p, 0, n * sizeof(T), st));
  // z = P r and partials with alpha = 0 (no x/r change)
  dgc::CgParams<T> cp{};
  cp.K = s->disc->K;
  cp.x = L.x;
  cp.r = L.r;
  cp.p = L.p;
  cp.q = L.q;
  cp.z = L.z;
  cp.P = L.use_P ? L.P : nullptr;
  cp.scal = init_scal;
  cp.partial = L.partial;
  DGC_LAUNCH(dgc::cg_update_launch<T>(s->disc->order, cp, st));
  if constexpr (std::is_same<T, double>::value) {
    // the done flag is set while initialising: run the correction unconditionally
    if (L.use_coarse) {
      int rc = coarse_correct(s, L.r, L.z, L.partial, 0, nullptr);
      if (rc) return rc;
    }
  }
  if (!fused) DGC_LAUNCH(dgc::cg_scalars_launch<T>(L.partial, s->grid, dgc::kInit, L.scal, history, st));
  return DGC_OK;
}

// One PCG iteration. Pass 1 forms the search direction p' = z + beta p (into
// the other direction buffer) and applies the first operator pass to it;
// pass 2 gives q = S p' and the partials of p'.q; the update forms x, r and
// z = P r with the partials (r.z, r.r) [two-level: then the coarse
// restriction, solve and prolongation, which rewrites the partials].
// Unfused, the scalar kernel finishes alpha after pass 2 and beta, the history
// and the stopping test after the update (5 kernels with block-Jacobi). Fused,
// the update sums pass 2's partials itself for alpha, and the next
// iteration's pass 1 sums the update's partials for beta and the stopping
// test, reading record it % 2 and writing the other (3 kernels).
// Iteration `it` of a solve reads direction buffer it % 2 and writes the other.
template <typename T>
static int cg_iteration(dgc_ipdg_s* s, IpdgLevel<T>& L, double* history, int it) {
  hipStream_t st = s->disc->ctx->stream;
  const bool fused = fused_iteration(s);
  T* pold = (it & 1) ? L.p2 : L.p;
  T* pnew = (it & 1) ? L.p : L.p2;
  T* rin = L.rec + (it & 1) * dgc::kRecStride;
  T* rout = L.rec + ((it + 1) & 1) * dgc::kRecStride;
  T* scal = fused ? rout : L.scal;  // what the kernels after pass 1 read
  dgc::IpdgParams<T> pr = L.params(*s);
  pr.u = pold;
  pr.dirz = L.z;
  pr.dirp = pnew;
  pr.out = L.q;
  pr.dotv = pnew;
  if (fused) {
    pr.scal = rin;
    pr.rec_out = rout;
    pr.upart = L.partial;
    pr.nparts = s->grid;
    pr.history = history;
    DGC_LAUNCH(dgc::ipdg_apply<T>(s->disc->order, pr, dgc::kPhaseOne, st));
    pr.scal = rout;
    pr.partial = L.dpartial;
    DGC_LAUNCH(dgc::ipdg_apply<T>(s->disc->order, pr, dgc::kSpdPhaseTwo, st));
  } else {
    pr.scal = L.scal;
    pr.partial = L.partial;
    DGC_LAUNCH(dgc::ipdg_apply<T>(s->disc->order, pr, dgc::kSpd, st));
    DGC_LAUNCH(dgc::cg_scalars_launch<T>(L.partial, s->grid, dgc::kAfterApply, L.scal, history, st));
  }
  dgc::CgParams<T> cp{};
  cp.K = s->disc->K;
  cp.x = L.x;
  cp.r = L.r;
  cp.p = pnew;
  cp.q = L.q;
  cp.z = L.z;
  cp.P = L.use_P ? L.P : nullptr;
  cp.scal = scal;
  cp.partial = L.partial;
  if (fused) {
    cp.dpart = L.dpartial;
    cp.nparts = s->grid;
    cp.rec_out = rout;
  }
  DGC_LAUNCH(dgc::cg_update_launch<T>(s->disc->order, cp, st));
  if constexpr (std::is_same<T, double>::value) {
    if (L.use_coarse) {
      int rc = coarse_correct(s, L.r, L.z, L.partial, 0, scal);
      if (rc) return rc;
    }
  }
  if (!fused)
    DGC_LAUNCH(dgc::cg_scalars_launch<T>(L.partial, s->grid, dgc::kAfterUpdate, L.scal, history, st));
  return DGC_OK;
}

// Run iterations in graph-captured chunks until the device flags completion.
template <typename T>
static int cg_run(dgc_ipdg_s* s, IpdgLevel<T>& L, double* history, int maxit, int* iters) {
  hipStream_t st = s->disc->ctx->stream;
  const int chunk = 16;  // even: every chunk starts on direction buffer 0
  if (s->use_graphs &&
      (!L.graph || L.graph_hist != history || L.graph_P != L.use_P || L.graph_coarse != L.use_coarse)) {
    if (L.graph) (void)hipGraphExecDestroy(L.graph);
    L.graph = nullptr;
    hipGraph_t g = nullptr;
    DGC_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = DGC_OK;
    for (int i = 0; i < chunk && rc == DGC_OK; ++i) rc = cg_iteration(s, L, history, i);
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != DGC_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return dgc_set_error(DGC_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&L.graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      L.graph = nullptr;
      return dgc_set_error(DGC_ERR_HIP, std::string("instantiate: ") + hipGetErrorString(e));
    }
    L.graph_hist = history;
    L.graph_P = L.use_P;
    L.graph_coarse = L.use_coarse;
  }
  // after whole chunks the current scalars are in record 0 (fused) or L.scal
  const T* cur = fused_iteration(s) ? L.rec : L.scal;
  T h[dgc::kCgScalars];
  const bool snapping = std::is_same<T, double>::value && s->snap != nullptr;
  const size_t n = size_t(s->disc->K) * s->disc->Np;
  bool snapped = false;  // an iterate was recorded after the last replay
  for (;;) {
    DGC_HIP_TRY(hipMemcpyAsync(h, cur, sizeof(h), hipMemcpyDeviceToHost, st));
    DGC_HIP_TRY(hipStreamSynchronize(st));
    if (h[dgc::kDone] != T(0)) {
      if (snapped) --s->snap_taken;  // recorded after (or at) convergence
      break;
    }
    if (int(h[dgc::kIter]) >= maxit) break;
    if (s->use_graphs) {
      DGC_HIP_TRY(hipGraphLaunch(L.graph, st));
    } else {
      for (int i = 0; i < chunk; ++i) {
        int rc = cg_iteration(s, L, history, i);
        if (rc) return rc;
      }
    }
    snapped = false;
    if (snapping && s->snap_taken < s->nsnap) {
      DGC_HIP_TRY(hipMemcpyAsync(s->snap + size_t(s->snap_taken) * n, L.x, n * sizeof(double),
                                 hipMemcpyDeviceToDevice, st));
      ++s->snap_taken;
      snapped = true;
    }
  }
  *iters = int(h[dgc::kIter]);
  return DGC_OK;
}

// ||b - S x||^2 on the device (S x into q, b - S x into r); waits for the value.
int true_residual2(dgc_ipdg_s* s, const double* b, const double* x, double* q, double* r, double* rr) {
  hipStream_t st = s->disc->ctx->stream;
  const size_t n = size_t(s->disc->K) * s->disc->Np;
  dgc::IpdgParams<double> pr = s->f64.params(*s);
  pr.u = x;
  pr.out = q;
  DGC_LAUNCH(dgc::ipdg_apply<double>(s->disc->order, pr, dgc::kSpd, st));
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(256), 0, st, 1.0, b, -1.0, (const double*)q, r, n);
  DGC_LAUNCH(hipGetLastError());
  return dot2(s, r, r, rr);
}

// v.w on the device; waits for the value.
int dot2(dgc_ipdg_s* s, const double* v, const double* w, double* out) {
  hipStream_t st = s->disc->ctx->stream;
  const size_t n = size_t(s->disc->K) * s->disc->Np;
  DGC_LAUNCH(dgc::dot_device(v, w, n, s->dot, s->dot + 256, st));
  DGC_HIP_TRY(hipMemcpyAsync(out, s->dot + 256, sizeof(double), hipMemcpyDeviceToHost, st));
  DGC_HIP_TRY(hipStreamSynchronize(st));
  return DGC_OK;
}

// FP64 PCG restart from the current iterate (used by the flexible-CG fallback).
int cg_init_fp64_restart(dgc_ipdg_s* s, const double* b, double tol2_times_bb, int maxit, double* history) {
  return cg_init<double>(s, s->f64, b, tol2_times_bb, maxit, false, history);
}

int cg_run_fp64(dgc_ipdg_s* s, double* history, int maxit, int* iters) {
  return cg_run<double>(s, s->f64, history, maxit, iters);
}

static void release_coarse(dgc_ipdg_s* s);

// Everything derived from the SPD operator is stale after a change of its
// boundary kind or shift: block inverses, coarse space, captured solver
// graphs (they hold the old parameters).
static void operator_changed(dgc_ipdg_s* s) {
  ++s->op_gen;
  release_coarse(s);
  (void)hipFree(s->f64.P);
  s->f64.P = nullptr;
  s->f64.use_P = false;
  (void)hipFree(s->f32.P);
  s->f32.P = nullptr;
  s->f32.use_P = false;
  if (s->f32.graph) (void)hipGraphExecDestroy(s->f32.graph);
  s->f32.graph = nullptr;
  s->diag_blocks.clear();
}

static void face_penalty(const dgc_disc_s* d, int tau_mode, double tau_scale, std::vector<double>& tau) {
  const int K = d->K, N = d->order;
  tau.assign(3 * size_t(K), tau_scale);
  if (tau_mode == 0) return;
  for (int c = 0; c < K; ++c)
    for (int e = 0; e < 3; ++e) {
      double h = 2.0 * d->h_J[c] / d->h_face_J[3 * size_t(c) + e];
      const int code = d->h_nbr[size_t(e) * K + c];
      if ((code & 3) != dgc::kBoundaryCode) {
        const int nc = code >> 2, nf = code & 3;
        h = std::min(h, 2.0 * d->h_J[nc] / d->h_face_J[3 * size_t(nc) + nf]);
      }
      tau[size_t(e) * K + c] = tau_scale * (N + 1.0) * (N + 2.0) / h;
    }
}

extern "C" {

int dgc_ipdg_create(dgc_disc_t disc, const double* Dr, const double* Ds, const double* lift,
                    const double* mass, int tau_mode, double tau_scale, dgc_ipdg_t* out) {
  DGC_REQUIRE(disc != nullptr && out != nullptr, "dgc_ipdg_create: null argument");
  DGC_REQUIRE(Dr && Ds && lift && mass, "dgc_ipdg_create: null operator matrix");
  DGC_REQUIRE(tau_mode == 0 || tau_mode == 1, "dgc_ipdg_create: tau_mode must be 0 or 1");
  DGC_REQUIRE(std::isfinite(tau_scale) && tau_scale > 0, "dgc_ipdg_create: tau_scale must be positive");
  DGC_HIP_TRY(hipSetDevice(disc->ctx->device));
  auto* s = new dgc_ipdg_s;
  s->disc = disc;
  s->tau_mode = tau_mode;
  s->tau_scale = tau_scale;
  s->grid = (disc->K + dgc::kTile - 1) / dgc::kTile;
  const int Np = disc->Np, NF = 3 * disc->Nfp, K = disc->K;
  s->mass.assign(mass, mass + size_t(Np) * Np);
  std::vector<double> tau;
  face_penalty(disc, tau_mode, tau_scale, tau);
  std::vector<double> geo(size_t(dgc::kGeoComps) * K), J(disc->h_J);
  int rc = DGC_OK;
  if (hipMemcpy(geo.data(), disc->geo, geo.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    rc = dgc_set_error(DGC_ERR_HIP, "dgc_ipdg_create: geometry readback failed");
  // the SPD form's tables M Dr, M Ds, M LIFT (row-major products)
  auto mass_times = [&](const double* A, int cols) {
    std::vector<double> P(size_t(Np) * cols, 0.0);
    for (int i = 0; i < Np; ++i)
      for (int k = 0; k < Np; ++k)
        for (int j = 0; j < cols; ++j) P[size_t(i) * cols + j] += mass[size_t(i) * Np + k] * A[size_t(k) * cols + j];
    return P;
  };
  const auto MDr = mass_times(Dr, Np), MDs = mass_times(Ds, Np), ML = mass_times(lift, NF);
  const std::vector<std::vector<double>> tables = {
      transpose_pad(Dr, Np, Np),         transpose_pad(Ds, Np, Np),         transpose_pad(lift, Np, NF),
      transpose_pad(mass, Np, Np),       transpose_pad(MDr.data(), Np, Np), transpose_pad(MDs.data(), Np, Np),
      transpose_pad(ML.data(), Np, NF)};
  if (rc == DGC_OK) rc = s->f64.init(*s, tables, geo, J, tau);
  if (rc == DGC_OK) rc = s->f32.init(*s, tables, geo, J, tau);
  if (rc == DGC_OK) rc = alloc(&s->hist, 1);
  if (rc == DGC_OK) rc = alloc(&s->dot, 257);
  if (rc != DGC_OK) {
    s->f64.release();
    s->f32.release();
    (void)hipFree(s->hist);
    (void)hipFree(s->dot);
    delete s;
    return rc;
  }
  *out = s;
  return DGC_OK;
}

int dgc_ipdg_destroy(dgc_ipdg_t s) {
  if (!s) return DGC_OK;
  (void)hipStreamSynchronize(s->disc->ctx->stream);
  release_coarse(s);
  s->f64.release();
  s->f32.release();
  (void)hipFree(s->hist);
  (void)hipFree(s->dot);
  (void)hipFree(s->color);
  delete s;
  return DGC_OK;
}

int dgc_ipdg_apply(dgc_ipdg_t s, double* u) {
  DGC_REQUIRE(s && u, "dgc_ipdg_apply: null argument");
  dgc::IpdgParams<double> pr = s->f64.params(*s);
  pr.u = u;
  pr.out = u;  // pass 2 reads only q and du, so the in-place update is safe
  DGC_LAUNCH(dgc::ipdg_apply<double>(s->disc->order, pr, dgc::kLaplace, s->disc->ctx->stream));
  return DGC_OK;
}

int dgc_ipdg_phase(dgc_ipdg_t s, int phase, double* u) {
  DGC_REQUIRE(s && u, "dgc_ipdg_phase: null argument");
  DGC_REQUIRE(phase == 1 || phase == 2, "dgc_ipdg_phase: phase must be 1 or 2");
  dgc::IpdgParams<double> pr = s->f64.params(*s);
  pr.u = u;
  pr.out = u;
  DGC_LAUNCH(dgc::ipdg_apply<double>(s->disc->order, pr, phase == 1 ? dgc::kPhaseOne : dgc::kPhaseTwo,
                                     s->disc->ctx->stream));
  return DGC_OK;
}

int dgc_ipdg_apply_spd(dgc_ipdg_t s, const double* u, double* out) {
  DGC_REQUIRE(s && u && out && u != out, "dgc_ipdg_apply_spd: bad argument");
  dgc::IpdgParams<double> pr = s->f64.params(*s);
  pr.u = u;
  pr.out = out;
  DGC_LAUNCH(dgc::ipdg_apply<double>(s->disc->order, pr, dgc::kSpd, s->disc->ctx->stream));
  return DGC_OK;
}

int dgc_ipdg_set_boundary(dgc_ipdg_t s, int kind) {
  DGC_REQUIRE(s != nullptr, "dgc_ipdg_set_boundary: null solver");
  DGC_REQUIRE(kind == DGC_BC_DIRICHLET || kind == DGC_BC_NEUMANN, "dgc_ipdg_set_boundary: unknown kind");
  if (kind == s->neumann) return DGC_OK;
  DGC_HIP_TRY(hipStreamSynchronize(s->disc->ctx->stream));
  s->neumann = kind;
  operator_changed(s);
  return DGC_OK;
}

int dgc_ipdg_set_graphs(dgc_ipdg_t s, int enabled) {
  DGC_REQUIRE(s != nullptr, "dgc_ipdg_set_graphs: null solver");
  s->use_graphs = enabled != 0;
  return DGC_OK;
}

int dgc_ipdg_set_shift(dgc_ipdg_t s, double sigma) {
  DGC_REQUIRE(s != nullptr, "dgc_ipdg_set_shift: null solver");
  DGC_REQUIRE(std::isfinite(sigma) && sigma >= 0.0, "dgc_ipdg_set_shift: need a finite sigma >= 0");
  if (sigma == s->shift) return DGC_OK;
  DGC_HIP_TRY(hipStreamSynchronize(s->disc->ctx->stream));
  s->shift = sigma;
  operator_changed(s);
  return DGC_OK;
}

int dgc_ipdg_apply_data(dgc_ipdg_t s, const double* gd, const double* hd, double* u) {
  DGC_REQUIRE(s && u, "dgc_ipdg_apply_data: null argument");
  dgc::IpdgParams<double> pr = s->f64.params(*s);
  pr.u = u;
  pr.out = u;
  pr.gbd = gd;
  pr.hbd = hd;
  DGC_LAUNCH(dgc::ipdg_apply<double>(s->disc->order, pr, dgc::kLaplace, s->disc->ctx->stream));
  return DGC_OK;
}

int dgc_ipdg_apply_inhomogeneous(dgc_ipdg_t s, const double* gu, const double* hq, double* inout) {
  DGC_REQUIRE(s && inout, "dgc_ipdg_apply_inhomogeneous: null argument");
  dgc::IpdgParams<double> pr = s->f64.params(*s);
  DGC_LAUNCH(dgc::ipdg_inhomogeneous(s->disc->order, pr, gu, hq, inout, s->disc->ctx->stream));
  return DGC_OK;
}

int dgc_ipdg_mass(dgc_ipdg_t s, const double* f, double* out) {
  DGC_REQUIRE(s && f && out && f != out, "dgc_ipdg_mass: bad argument");
  dgc::IpdgParams<double> pr = s->f64.params(*s);
  pr.u = f;
  pr.out = out;
  DGC_LAUNCH(dgc::ipdg_mass_apply<double>(s->disc->order, pr, s->disc->ctx->stream));
  return DGC_OK;
}

int dgc_ipdg_block_jacobi(dgc_ipdg_t s, int* num_colors) {
  DGC_REQUIRE(s != nullptr, "dgc_ipdg_block_jacobi: null solver");
  dgc_disc_s* d = s->disc;
  const int K = d->K, Np = d->Np;
  hipStream_t st = d->ctx->stream;
  // Distance-2 greedy colouring: (S u)_c depends on u in c and in cells up to
  // two faces away, so cells of one colour can be probed together.
  std::vector<int> nb(3 * size_t(K), -1);
  for (int c = 0; c < K; ++c)
    for (int e = 0; e < 3; ++e) {
      const int code = d->h_nbr[size_t(e) * K + c];
      if ((code & 3) != dgc::kBoundaryCode) nb[3 * size_t(c) + e] = code >> 2;
    }
  std::vector<int> color(K, -1), mark(64, -1);
  int ncol = 0;
  for (int c = 0; c < K; ++c) {
    std::vector<int> near;
    for (int e = 0; e < 3; ++e) {
      const int a = nb[3 * size_t(c) + e];
      if (a < 0) continue;
      near.push_back(a);
      for (int f = 0; f < 3; ++f) {
        const int b = nb[3 * size_t(a) + f];
        if (b >= 0 && b != c) near.push_back(b);
      }
    }
    // also cells two faces away through either side (symmetric closure)
    for (int a : near)
      if (color[a] >= 0 && color[a] < 64) mark[color[a]] = c;
    int col = 0;
    while (col < 64 && mark[col] == c) ++col;
    if (col >= 64) return dgc_set_error(DGC_ERR_UNSUPPORTED, "block-Jacobi colouring needs > 64 colours");
    color[c] = col;
    ncol = std::max(ncol, col + 1);
  }
  int rc;
  if (!s->color && (rc = alloc(&s->color, K))) return rc;
  DGC_HIP_TRY(hipMemcpy(s->color, color.data(), K * sizeof(int), hipMemcpyHostToDevice));
  const size_t n = size_t(K) * Np;
  double *u = s->f64.z, *o = s->f64.q, *blocks = nullptr;
  if ((rc = alloc(&blocks, size_t(Np) * Np * K))) return rc;
  for (int col = 0; col < ncol; ++col)
    for (int k = 0; k < Np; ++k) {
      hipLaunchKernelGGL(probe_fill_kernel, dim3(grid_for(n)), dim3(256), 0, st, s->color, col, k, Np, K, u);
      DGC_LAUNCH(hipGetLastError());
      dgc::IpdgParams<double> pr = s->f64.params(*s);
      pr.u = u;
      pr.out = o;
      DGC_LAUNCH(dgc::ipdg_apply<double>(d->order, pr, dgc::kSpd, st));
      hipLaunchKernelGGL(probe_extract_kernel, dim3((K + 255) / 256), dim3(256), 0, st, s->color, col, k, Np, K,
                         o, blocks);
      DGC_LAUNCH(hipGetLastError());
    }
  std::vector<double> hb(size_t(Np) * Np * K), hinv(size_t(Np) * (Np + 1) / 2 * K);
  hipError_t e = hipMemcpyAsync(hb.data(), blocks, hb.size() * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(blocks);
  if (e != hipSuccess) return dgc_set_error(DGC_ERR_HIP, "block readback failed");
  dgc_host::Mat B(Np, Np);
  for (int c = 0; c < K; ++c) {
    for (int j = 0; j < Np; ++j)
      for (int k = 0; k < Np; ++k) B(j, k) = hb[(size_t(j) * Np + k) * K + c];
    dgc_host::Mat Bi;
    try {
      Bi = dgc_host::inverse(B);
    } catch (const std::exception& ex) {
      return dgc_set_error(DGC_ERR_INVALID, std::string("block-Jacobi: cell block not invertible: ") + ex.what());
    }
    // symmetric by construction: store the mean of (j,k) and (k,j), upper triangle
    for (int j = 0; j < Np; ++j)
      for (int k = j; k < Np; ++k)
        hinv[size_t(dgc::packed_index(Np, j, k)) * K + c] = 0.5 * (Bi(j, k) + Bi(k, j));
  }
  // captured solver graphs (PCG at both precisions, the network's FCG) hold
  // the old inverses' addresses
  if (s->f64.graph) (void)hipGraphExecDestroy(s->f64.graph);
  s->f64.graph = nullptr;
  if (s->f32.graph) (void)hipGraphExecDestroy(s->f32.graph);
  s->f32.graph = nullptr;
  ++s->coarse_gen;
  (void)hipFree(s->f64.P);
  (void)hipFree(s->f32.P);
  s->f64.P = nullptr;
  s->f32.P = nullptr;
  if ((rc = upload(&s->f64.P, hinv)) || (rc = upload(&s->f32.P, cast<float>(hinv)))) return rc;
  s->diag_blocks = hb;
  if (num_colors) *num_colors = ncol;
  return DGC_OK;
}

int dgc_ipdg_diag_blocks(dgc_ipdg_t s, double* blocks) {
  DGC_REQUIRE(s && blocks, "dgc_ipdg_diag_blocks: null argument");
  DGC_REQUIRE(!s->diag_blocks.empty(), "dgc_ipdg_diag_blocks: call dgc_ipdg_block_jacobi first");
  std::copy(s->diag_blocks.begin(), s->diag_blocks.end(), blocks);
  return DGC_OK;
}

}  // extern "C"

static void release_coarse(dgc_ipdg_s* s) {
  // captured solver graphs hold the coarse arrays' addresses
  if (s->f64.graph) (void)hipGraphExecDestroy(s->f64.graph);
  s->f64.graph = nullptr;
  ++s->coarse_gen;
  for (void* p : s->coarse_owned) (void)hipFree(p);
  s->coarse_owned.clear();
  s->coarse = dgc::CoarseSpace{};
  s->coarse_A.clear();
  s->f64.use_coarse = false;
}

template <typename T>
static int coarse_alloc(dgc_ipdg_s* s, T** p, size_t n) {
  int rc = alloc(p, n);
  if (rc == DGC_OK) s->coarse_owned.push_back(*p);
  return rc;
}

int coarse_correct(dgc_ipdg_s* s, const double* r, double* z, double* partial, int zz, const double* scal) {
  hipStream_t st = s->disc->ctx->stream;
  const int N = s->disc->order;
  DGC_LAUNCH(dgc::coarse_restrict(N, s->coarse, r, s->coarse.rc, 1, scal, st));
  DGC_LAUNCH(dgc::coarse_solve(s->coarse, scal, st));
  DGC_LAUNCH(dgc::coarse_prolong(N, s->coarse, r, z, partial, zz, scal, st));
  return DGC_OK;
}

extern "C" {

int dgc_ipdg_coarse_space(dgc_ipdg_t s, const double* cell_x, const double* cell_y, int gx, int gy,
                          int* num_coarse) {
  DGC_REQUIRE(s && cell_x && cell_y, "dgc_ipdg_coarse_space: null argument");
  DGC_REQUIRE(gx >= 1 && gy >= 1, "dgc_ipdg_coarse_space: need gx, gy >= 1");
  dgc_disc_s* d = s->disc;
  const int K = d->K, Np = d->Np, N = d->order;
  const size_t n = size_t(K) * Np;
  hipStream_t st = d->ctx->stream;
  DGC_HIP_TRY(hipStreamSynchronize(st));
  release_coarse(s);
  // aggregates: the cells whose centroid falls in one box of a gx x gy grid
  // over the bounding box of the vertices; empty boxes are dropped
  double x0 = cell_x[0], x1 = cell_x[0], y0 = cell_y[0], y1 = cell_y[0];
  for (size_t i = 0; i < 3 * size_t(K); ++i) {
    DGC_REQUIRE(std::isfinite(cell_x[i]) && std::isfinite(cell_y[i]), "dgc_ipdg_coarse_space: bad vertex");
    x0 = std::min(x0, cell_x[i]);
    x1 = std::max(x1, cell_x[i]);
    y0 = std::min(y0, cell_y[i]);
    y1 = std::max(y1, cell_y[i]);
  }
  const double wx = (x1 - x0) / gx, wy = (y1 - y0) / gy;
  DGC_REQUIRE(wx > 0 && wy > 0, "dgc_ipdg_coarse_space: degenerate mesh extent");
  std::vector<int> box(K), box_agg(size_t(gx) * gy, -1);
  for (int c = 0; c < K; ++c) {
    const double cx = (cell_x[3 * c] + cell_x[3 * c + 1] + cell_x[3 * c + 2]) / 3.0;
    const double cy = (cell_y[3 * c] + cell_y[3 * c + 1] + cell_y[3 * c + 2]) / 3.0;
    const int ix = std::min(gx - 1, std::max(0, int((cx - x0) / wx)));
    const int iy = std::min(gy - 1, std::max(0, int((cy - y0) / wy)));
    box[c] = ix + gx * iy;
    box_agg[box[c]] = 0;
  }
  int na = 0;
  for (auto& b : box_agg)
    if (b == 0) b = na++;
  const int nc = 3 * na, ld = (nc + 3) / 4 * 4;
  DGC_REQUIRE(nc <= 6144, "dgc_ipdg_coarse_space: more than 2048 aggregates (use a coarser grid)");
  std::vector<int> cell_agg(K), ptr(na + 1, 0), cells(K);
  for (int c = 0; c < K; ++c) ptr[(cell_agg[c] = box_agg[box[c]]) + 1]++;
  for (int a = 0; a < na; ++a) ptr[a + 1] += ptr[a];
  {
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int c = 0; c < K; ++c) cells[fill[cell_agg[c]]++] = c;
  }
  // basis coefficients: x(rho, sigma) = (v2 + v3)/2 + rho (v2 - v1)/2 + sigma (v3 - v1)/2
  std::vector<float> coef(6 * size_t(K));
  const double h = 0.5 * std::max(wx, wy);
  for (int c = 0; c < K; ++c) {
    const int b = box[c], ix = b % gx, iy = b / gx;
    const double xa = x0 + (ix + 0.5) * wx, ya = y0 + (iy + 0.5) * wy;
    const double* vx = cell_x + 3 * size_t(c);
    const double* vy = cell_y + 3 * size_t(c);
    const double t[6] = {(0.5 * (vx[1] + vx[2]) - xa) / h, 0.5 * (vx[1] - vx[0]) / h, 0.5 * (vx[2] - vx[0]) / h,
                         (0.5 * (vy[1] + vy[2]) - ya) / h, 0.5 * (vy[1] - vy[0]) / h, 0.5 * (vy[2] - vy[0]) / h};
    for (int k = 0; k < 6; ++k) coef[size_t(k) * K + c] = float(t[k]);
  }
  const dgc_host::RefTriangle ref(N, dgc_host::kStrongForm);
  std::vector<double> rs(2 * size_t(Np));
  for (int j = 0; j < Np; ++j) {
    rs[j] = ref.r[j];
    rs[Np + j] = ref.s[j];
  }
  dgc::CoarseSpace& cs = s->coarse;
  cs.K = K;
  cs.na = na;
  cs.nc = nc;
  cs.ld = ld;
  int *dptr = nullptr, *dcells = nullptr, *dagg = nullptr;
  float *dcoef = nullptr, *dAinv = nullptr;
  double* drs = nullptr;
  int rc = DGC_OK;
  if ((rc = coarse_alloc(s, &dptr, ptr.size())) || (rc = coarse_alloc(s, &dcells, size_t(K))) ||
      (rc = coarse_alloc(s, &dagg, size_t(K))) || (rc = coarse_alloc(s, &dcoef, coef.size())) ||
      (rc = coarse_alloc(s, &drs, rs.size())) || (rc = coarse_alloc(s, &dAinv, size_t(nc) * ld)) ||
      (rc = coarse_alloc(s, &cs.rc, size_t(ld))) || (rc = coarse_alloc(s, &cs.yc, size_t(ld)))) {
    release_coarse(s);
    return rc;
  }
  // the solve reads rc in 4-entry groups up to ld: the padding must be zero
  DGC_HIP_TRY(hipMemset(cs.rc, 0, size_t(ld) * sizeof(double)));
  DGC_HIP_TRY(hipMemcpy(dptr, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice));
  DGC_HIP_TRY(hipMemcpy(dcells, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice));
  DGC_HIP_TRY(hipMemcpy(dagg, cell_agg.data(), cell_agg.size() * sizeof(int), hipMemcpyHostToDevice));
  DGC_HIP_TRY(hipMemcpy(dcoef, coef.data(), coef.size() * sizeof(float), hipMemcpyHostToDevice));
  DGC_HIP_TRY(hipMemcpy(drs, rs.data(), rs.size() * sizeof(double), hipMemcpyHostToDevice));
  cs.agg_ptr = dptr;
  cs.agg_cells = dcells;
  cs.cell_agg = dagg;
  cs.coef = dcoef;
  cs.rs = drs;
  cs.Ainv = dAinv;
  // Galerkin matrix column by column: A[:, col] = Phi^T S phi_col
  double *A = nullptr, *u = nullptr, *w = nullptr, *work = nullptr;
  auto cleanup = [&]() {
    (void)hipFree(A);
    (void)hipFree(u);
    (void)hipFree(w);
    (void)hipFree(work);
  };
  if ((rc = alloc(&A, size_t(nc) * nc)) || (rc = alloc(&u, n)) || (rc = alloc(&w, n)) ||
      (rc = alloc(&work, 2 * size_t(nc)))) {
    cleanup();
    release_coarse(s);
    return rc;
  }
  dgc::IpdgParams<double> pr = s->f64.params(*s);
  pr.u = u;
  pr.out = w;
  hipError_t e = hipSuccess;
  for (int col = 0; col < nc && e == hipSuccess; ++col) {
    e = hipMemsetAsync(u, 0, n * sizeof(double), st);
    if (e == hipSuccess) e = dgc::coarse_fill(N, cs, col, u, st);
    if (e == hipSuccess) e = dgc::ipdg_apply<double>(N, pr, dgc::kSpd, st);
    if (e == hipSuccess) e = dgc::coarse_restrict(N, cs, w, A + col, nc, nullptr, st);
  }
  if (e == hipSuccess) {
    s->coarse_A.resize(size_t(nc) * nc);
    e = hipMemcpyAsync(s->coarse_A.data(), A, s->coarse_A.size() * sizeof(double), hipMemcpyDeviceToHost, st);
  }
  if (e == hipSuccess && s->neumann && s->shift == 0.0) {
    // With Neumann faces everywhere (and no shift) the constants are in the null space of S,
    // and so is their coarse representation z (Phi z = 1, exact: every
    // aggregate carries the constants). Solve for z aggregate by aggregate
    // from the Gram blocks Phi_a^T Phi_a z_a = Phi_a^T 1, then invert
    // A + alpha z z^T: for the PCG residuals, orthogonal to the constants,
    // the coarse right-hand side is orthogonal to z and the correction is
    // the one of A's pseudo-inverse.
    std::vector<double> G(size_t(nc) * 3), rhs(nc);  // block diagonal: 3 entries per column
    e = hipMemsetAsync(u, 0, n * sizeof(double), st);
    std::vector<double> ones(n, 1.0);
    if (e == hipSuccess) e = hipMemcpyAsync(u, ones.data(), n * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = dgc::coarse_restrict(N, cs, u, A, 1, nullptr, st);  // Phi^T 1 into A's first row
    std::vector<double> col(size_t(nc) * nc);
    if (e == hipSuccess) e = hipMemcpyAsync(rhs.data(), A, nc * sizeof(double), hipMemcpyDeviceToHost, st);
    for (int c = 0; c < nc && e == hipSuccess; ++c) {
      e = hipMemsetAsync(u, 0, n * sizeof(double), st);
      if (e == hipSuccess) e = dgc::coarse_fill(N, cs, c, u, st);
      if (e == hipSuccess) e = dgc::coarse_restrict(N, cs, u, A + c, nc, nullptr, st);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(col.data(), A, col.size() * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    std::vector<double> z(nc, 0.0);
    for (int a = 0; e == hipSuccess && a < nc / 3; ++a) {
      double g[3][4];
      for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) g[i][j] = col[size_t(3 * a + i) * nc + (3 * a + j)];
        g[i][3] = rhs[3 * a + i];
      }
      for (int i = 0; i < 3; ++i) {  // Gauss-Jordan with partial pivoting on the 3 x 3 block
        int piv = i;
        for (int j = i + 1; j < 3; ++j)
          if (std::abs(g[j][i]) > std::abs(g[piv][i])) piv = j;
        for (int k = 0; k < 4; ++k) std::swap(g[i][k], g[piv][k]);
        const double d = g[i][i];
        for (int k = 0; k < 4; ++k) g[i][k] /= d;
        for (int j = 0; j < 3; ++j)
          if (j != i) {
            const double f = g[j][i];
            for (int k = 0; k < 4; ++k) g[j][k] -= f * g[i][k];
          }
      }
      for (int i = 0; i < 3; ++i) z[3 * a + i] = g[i][3];
    }
    double tr = 0.0, zz = 0.0;
    for (int i = 0; i < nc; ++i) {
      tr += s->coarse_A[size_t(i) * nc + i];
      zz += z[i] * z[i];
    }
    const double alpha = tr / nc / zz;
    std::vector<double> areg(s->coarse_A);
    for (int i = 0; i < nc; ++i)
      for (int j = 0; j < nc; ++j) areg[size_t(i) * nc + j] += alpha * z[i] * z[j];
    if (e == hipSuccess) e = hipMemcpyAsync(A, areg.data(), areg.size() * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  if (e == hipSuccess) e = dgc::spd_inverse_inplace(A, nc, work, st);
  if (e == hipSuccess) e = dgc::symmetrize_to_float(A, nc, ld, dAinv, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  cleanup();
  if (e != hipSuccess) {
    release_coarse(s);
    return dgc_set_error(DGC_ERR_HIP, std::string("dgc_ipdg_coarse_space: ") + hipGetErrorString(e));
  }
  for (int i = 0; i < nc; ++i)
    if (!(s->coarse_A[size_t(i) * nc + i] > 0)) {
      release_coarse(s);
      return dgc_set_error(DGC_ERR_NONFINITE, "dgc_ipdg_coarse_space: coarse matrix is not positive definite");
    }
  std::vector<float> hinv(size_t(nc) * ld);
  DGC_HIP_TRY(hipMemcpy(hinv.data(), dAinv, hinv.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int i = 0; i < nc; ++i)
    if (!(hinv[size_t(i) * ld + i] > 0.f) || !std::isfinite(hinv[size_t(i) * ld + i])) {
      release_coarse(s);
      return dgc_set_error(DGC_ERR_NONFINITE, "dgc_ipdg_coarse_space: coarse inverse is not positive definite");
    }
  if (num_coarse) *num_coarse = nc;
  return DGC_OK;
}

int dgc_ipdg_coarse_info(dgc_ipdg_t s, double* A, float* Ainv, int* cell_agg, float* coef, double* rs) {
  DGC_REQUIRE(s, "dgc_ipdg_coarse_info: null solver");
  const dgc::CoarseSpace& cs = s->coarse;
  DGC_REQUIRE(cs.nc > 0, "dgc_ipdg_coarse_info: no coarse space");
  const int K = cs.K, nc = cs.nc;
  if (A) std::copy(s->coarse_A.begin(), s->coarse_A.end(), A);
  if (Ainv) DGC_HIP_TRY(hipMemcpy2D(Ainv, nc * sizeof(float), cs.Ainv, cs.ld * sizeof(float), nc * sizeof(float),
                                    nc, hipMemcpyDeviceToHost));
  if (cell_agg) DGC_HIP_TRY(hipMemcpy(cell_agg, cs.cell_agg, K * sizeof(int), hipMemcpyDeviceToHost));
  if (coef) DGC_HIP_TRY(hipMemcpy(coef, cs.coef, 6 * size_t(K) * sizeof(float), hipMemcpyDeviceToHost));
  if (rs) DGC_HIP_TRY(hipMemcpy(rs, cs.rs, 2 * size_t(s->disc->Np) * sizeof(double), hipMemcpyDeviceToHost));
  return DGC_OK;
}

int dgc_ipdg_pcg(dgc_ipdg_t s, const double* b, double* x, double rtol, int maxit, int precond,
                 int* iters, double* relres, double* history, float* ms) {
  DGC_REQUIRE(s && b && x, "dgc_ipdg_pcg: null argument");
  DGC_REQUIRE(rtol > 0 && maxit > 0, "dgc_ipdg_pcg: need rtol > 0 and maxit > 0");
  DGC_REQUIRE(precond >= 0 && precond <= 2,
              "dgc_ipdg_pcg: precond must be 0 (none), 1 (block-Jacobi) or 2 (two-level)");
  DGC_REQUIRE(precond != 2 || s->coarse.nc > 0, "dgc_ipdg_pcg: two-level needs dgc_ipdg_coarse_space first");
  if (precond >= 1 && !s->f64.P) {
    int rc = dgc_ipdg_block_jacobi(s, nullptr);
    if (rc) return rc;
  }
  dgc_disc_s* d = s->disc;
  hipStream_t st = d->ctx->stream;
  const size_t n = size_t(d->K) * d->Np;
  IpdgLevel<double>& L = s->f64;
  L.use_P = precond >= 1;
  L.use_coarse = precond == 2;
  if (s->hist_len < maxit + 1) {
    (void)hipFree(s->hist);
    s->hist = nullptr;
    int rc = alloc(&s->hist, size_t(maxit) + 1);
    if (rc) return rc;
    s->hist_len = maxit + 1;
    if (L.graph) (void)hipGraphExecDestroy(L.graph);
    L.graph = nullptr;
  }
  // ||b||^2 on the device (set-up, outside the timed region)
  double bb = 0.0;
  {
    int rc = dot2(s, b, b, &bb);
    if (rc) return rc;
  }
  DGC_REQUIRE(std::isfinite(bb), "dgc_ipdg_pcg: non-finite right-hand side");
  if (bb == 0.0) {
    DGC_HIP_TRY(hipMemsetAsync(x, 0, n * sizeof(double), st));
    if (iters) *iters = 0;
    if (relres) *relres = 0.0;
    return DGC_OK;
  }
  dgc::EventPair ev;
  DGC_HIP_TRY(ev.create());
  DGC_HIP_TRY(hipEventRecord(ev.start, st));
  DGC_HIP_TRY(hipMemsetAsync(L.x, 0, n * sizeof(double), st));
  int total = 0, rc = DGC_OK;
  double rel = 1.0;
  // restart from the true residual if recurrence drift left it above rtol
  for (int restart = 0; restart < 4 && rc == DGC_OK; ++restart) {
    rc = cg_init<double>(s, L, b, rtol * rtol * bb, maxit - total, restart == 0, s->hist + total);
    int it = 0;
    if (rc == DGC_OK) rc = cg_run<double>(s, L, s->hist + total, maxit - total, &it);
    s->snap = nullptr;  // restarts record nothing
    total += it;
    if (rc != DGC_OK) break;
    // true residual r = b - S x
    double rr = 0.0;
    if ((rc = true_residual2(s, b, L.x, L.q, L.r, &rr))) break;
    rel = std::sqrt(rr / bb);
    if (!(rel == rel)) {
      rc = dgc_set_error(DGC_ERR_NONFINITE, "dgc_ipdg_pcg: residual is not finite");
      break;
    }
    if (rel <= rtol || total >= maxit) break;
  }
  DGC_HIP_TRY(hipEventRecord(ev.stop, st));
  DGC_HIP_TRY(hipEventSynchronize(ev.stop));
  float elapsed = 0.f;
  (void)hipEventElapsedTime(&elapsed, ev.start, ev.stop);
  if (rc != DGC_OK) return rc;
  DGC_HIP_TRY(hipMemcpyAsync(x, L.x, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  if (history) DGC_HIP_TRY(hipMemcpyAsync(history, s->hist, (total + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
  DGC_HIP_TRY(hipStreamSynchronize(st));
  if (iters) *iters = total;
  if (relres) *relres = rel;
  if (ms) *ms = elapsed;
  if (rel > rtol) return dgc_set_error(DGC_ERR_NOT_CONVERGED, "dgc_ipdg_pcg: iteration cap reached");
  return DGC_OK;
}

int dgc_ipdg_pcg_mixed(dgc_ipdg_t s, const double* b, double* x, double rtol, double inner_rtol,
                       int max_outer, int max_inner, int* outer_iters, int* inner_iters, double* relres,
                       float* ms) {
  DGC_REQUIRE(s && b && x, "dgc_ipdg_pcg_mixed: null argument");
  DGC_REQUIRE(rtol > 0 && inner_rtol > 0 && inner_rtol < 1 && max_outer > 0 && max_inner > 0,
              "dgc_ipdg_pcg_mixed: bad tolerance or iteration cap");
  if (!s->f64.P) {
    int rc = dgc_ipdg_block_jacobi(s, nullptr);
    if (rc) return rc;
  }
  dgc_disc_s* d = s->disc;
  hipStream_t st = d->ctx->stream;
  const size_t n = size_t(d->K) * d->Np;
  IpdgLevel<double>& H = s->f64;
  IpdgLevel<float>& L = s->f32;
  L.use_P = true;
  double bb;
  int rc = dot2(s, b, b, &bb);
  if (rc) return rc;
  DGC_REQUIRE(std::isfinite(bb), "dgc_ipdg_pcg_mixed: non-finite right-hand side");
  dgc::EventPair ev;
  DGC_HIP_TRY(ev.create());
  DGC_HIP_TRY(hipEventRecord(ev.start, st));
  DGC_HIP_TRY(hipMemsetAsync(H.x, 0, n * sizeof(double), st));
  DGC_HIP_TRY(hipMemcpyAsync(H.r, b, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  double rr = bb, rel = bb > 0 ? 1.0 : 0.0;
  int outer = 0, inner_total = 0;
  if (s->hist_len < max_inner + 1) {
    (void)hipFree(s->hist);
    s->hist = nullptr;
    if ((rc = alloc(&s->hist, size_t(max_inner) + 1))) return rc;
    s->hist_len = max_inner + 1;
    if (L.graph) (void)hipGraphExecDestroy(L.graph);
    L.graph = nullptr;
    if (H.graph) (void)hipGraphExecDestroy(H.graph);
    H.graph = nullptr;
  }
  while (rel > rtol && outer < max_outer) {
    // inner FP32 solve of S e = r / ||r||
    const double rn = std::sqrt(rr);
    // The FP32 right-hand side goes to L.z: cg_init copies it into L.r before
    // anything overwrites z.
    hipLaunchKernelGGL(to_float_kernel, dim3(grid_for(n)), dim3(256), 0, st, (const double*)H.r, 1.0 / rn, L.z, n);
    DGC_LAUNCH(hipGetLastError());
    DGC_HIP_TRY(hipMemsetAsync(L.x, 0, n * sizeof(float), st));
    rc = cg_init<float>(s, L, L.z, inner_rtol * inner_rtol, max_inner, true, s->hist);
    int it = 0;
    if (rc == DGC_OK) rc = cg_run<float>(s, L, s->hist, max_inner, &it);
    if (rc) return rc;
    inner_total += it;
    // x += ||r|| e;  r = b - S x (FP64)
    hipLaunchKernelGGL(add_scaled_float_kernel, dim3(grid_for(n)), dim3(256), 0, st, H.x, rn, (const float*)L.x, n);
    DGC_LAUNCH(hipGetLastError());
    if ((rc = true_residual2(s, b, H.x, H.q, H.r, &rr))) return rc;
    rel = std::sqrt(rr / bb);
    ++outer;
    if (!(rel == rel)) return dgc_set_error(DGC_ERR_NONFINITE, "dgc_ipdg_pcg_mixed: residual is not finite");
  }
  DGC_HIP_TRY(hipEventRecord(ev.stop, st));
  DGC_HIP_TRY(hipEventSynchronize(ev.stop));
  float elapsed = 0.f;
  (void)hipEventElapsedTime(&elapsed, ev.start, ev.stop);
  DGC_HIP_TRY(hipMemcpyAsync(x, H.x, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  DGC_HIP_TRY(hipStreamSynchronize(st));
  if (outer_iters) *outer_iters = outer;
  if (inner_iters) *inner_iters = inner_total;
  if (relres) *relres = rel;
  if (ms) *ms = elapsed;
  if (rel > rtol) return dgc_set_error(DGC_ERR_NOT_CONVERGED, "dgc_ipdg_pcg_mixed: outer iteration cap reached");
  return DGC_OK;
}

}  // extern "C"
