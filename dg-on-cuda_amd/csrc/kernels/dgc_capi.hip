// C ABI over the gfx950 kernels: contexts, device memory, discretizations,
// semi-discrete operators and the RK4 time loop. See include/dgc_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/dgc_hip.h"
#include "dg_internal.hpp"

namespace {
thread_local std::string g_last_error;
}

// ---- work area of dgc_rk4_step_noalloc --------------------------------------
// A file-scope device array: every GPU that loads the library holds its own
// copy, at its own address, until the process ends. Host code never uses the
// array's name as a device address: dgc_ctx_create looks the address up on
// its device (hipGetSymbolAddress) and keeps it per device in g_scratch_ptr.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int dgc_set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define DGC_HIP_TRY(call)                                                                   \
  do {                                                                                      \
    hipError_t err_ = (call);                                                               \
    if (err_ != hipSuccess)                                                                 \
      return dgc_set_error(DGC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(err_)); \
  } while (0)

#define DGC_REQUIRE(cond, msg) \
  do {                         \
    if (!(cond)) return dgc_set_error(DGC_ERR_INVALID, msg); \
  } while (0)

// Full RK4 steps the time loops replay from one graph launch. On a mesh of a
// few tiles a step's four kernels run for a few microseconds each, and every
// graph launch adds its own dispatch and completion-signal round trip; a
// graph of 16 steps pays it once per 64 kernels.
constexpr int kStepsPerGraph = 16;

// A captured sequence of full steps, keyed on the buffers and dt it bakes in.
struct StepGraph {
  hipGraphExec_t exec = nullptr;
  std::vector<const void*> key;
  double dt = 0.0;
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key.clear();
  }
};

struct dgc_op_s {
  dgc_disc_s* disc = nullptr;
  int physics = 0;
  int F = 0;
  double* DrT = nullptr;
  double* DsT = nullptr;
  double* LT = nullptr;
  double ax = 0.0, ay = 0.0;
  int bc = 0;
  int wt = 4;  // waves per 64-cell tile of the stage kernels
  int precision = 0;      // 0: FP64; 1: FP32 matrix-core contractions (advection)
  float* opF = nullptr;   // matrix-core operand blocks of Dr, Ds, LIFT (advection)
  double* opD = nullptr;
  std::vector<double> hDr, hDs, hL;  // host copies of the operators
  int cache_policy = DGC_CACHE_AUTO;
  double* tclock = nullptr;
  int* flag = nullptr;  // non-finite flag of the time loops' state check
  // cached graphs of one full RK4 step and of kStepsPerGraph steps in a row
  StepGraph graph, chunk;
};

// Destroys the operator's cached step graphs (after a change of tiling,
// precision or cache policy, which the captured launches bake in).
static void drop_graphs(dgc_op_s* op) {
  op->graph.reset();
  op->chunk.reset();
}

extern "C" {

const char* dgc_last_error(void) { return g_last_error.c_str(); }

int dgc_max_order(void) { return dgc::kMaxOrder; }

int dgc_ctx_create(int device, dgc_ctx_t* out) {
  DGC_REQUIRE(out != nullptr, "dgc_ctx_create: out is null");
  int n = 0;
  DGC_HIP_TRY(hipGetDeviceCount(&n));
  DGC_REQUIRE(device >= 0 && device < n, "dgc_ctx_create: no such device");
  DGC_HIP_TRY(hipSetDevice(device));
  DGC_REQUIRE(device < kMaxDevices, "dgc_ctx_create: device index above 63");
  {
    // once per device, with that device current: the address of its work area
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device] == nullptr) {
      void* ptr = nullptr;
      DGC_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      g_scratch_ptr[device] = static_cast<double*>(ptr);
    }
  }
  auto* ctx = new dgc_ctx_s;
  ctx->device = device;
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return dgc_set_error(DGC_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  *out = ctx;
  return DGC_OK;
}

int dgc_ctx_destroy(dgc_ctx_t ctx) {
  if (ctx == nullptr) return DGC_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return DGC_OK;
}

int dgc_ctx_sync(dgc_ctx_t ctx) {
  DGC_REQUIRE(ctx != nullptr, "dgc_ctx_sync: null context");
  DGC_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return DGC_OK;
}

int dgc_malloc(dgc_ctx_t ctx, size_t bytes, void** ptr) {
  DGC_REQUIRE(ctx != nullptr && ptr != nullptr, "dgc_malloc: null argument");
  DGC_HIP_TRY(hipSetDevice(ctx->device));
  DGC_HIP_TRY(hipMalloc(ptr, bytes == 0 ? 1 : bytes));
  return DGC_OK;
}

int dgc_free(dgc_ctx_t ctx, void* ptr) {
  DGC_REQUIRE(ctx != nullptr, "dgc_free: null context");
  if (ptr == nullptr) return DGC_OK;
  DGC_HIP_TRY(hipStreamSynchronize(ctx->stream));
  DGC_HIP_TRY(hipFree(ptr));
  return DGC_OK;
}

int dgc_memcpy_h2d(dgc_ctx_t ctx, void* dst, const void* src, size_t bytes) {
  DGC_REQUIRE(ctx != nullptr, "dgc_memcpy_h2d: null context");
  DGC_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  DGC_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return DGC_OK;
}

int dgc_memcpy_d2h(dgc_ctx_t ctx, void* dst, const void* src, size_t bytes) {
  DGC_REQUIRE(ctx != nullptr, "dgc_memcpy_d2h: null context");
  DGC_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  DGC_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return DGC_OK;
}

int dgc_memcpy_d2d(dgc_ctx_t ctx, void* dst, const void* src, size_t bytes) {
  DGC_REQUIRE(ctx != nullptr, "dgc_memcpy_d2d: null context");
  DGC_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return DGC_OK;
}

int dgc_memset_zero(dgc_ctx_t ctx, void* dst, size_t bytes) {
  DGC_REQUIRE(ctx != nullptr, "dgc_memset_zero: null context");
  DGC_HIP_TRY(hipMemsetAsync(dst, 0, bytes, ctx->stream));
  return DGC_OK;
}

// ---------------------------------------------------------------------------
int dgc_disc2d_create(dgc_ctx_t ctx, int num_cells, int order, const int* face0_nodes,
                      const int* face1_nodes, const int* face2_nodes, const double* inv_jac,
                      const double* J, const double* face_J, const int* iface_cell,
                      const int* iface_face, int num_boundary_nodes, const double* bx,
                      const double* by, const double* nx, const double* ny, dgc_disc_t* out) {
  DGC_REQUIRE(ctx != nullptr && out != nullptr, "dgc_disc2d_create: null argument");
  DGC_REQUIRE(num_cells > 0 && num_cells < (1 << 29), "dgc_disc2d_create: num_cells out of range");
  if (order < 1 || order > dgc::kMaxOrder)
    return dgc_set_error(DGC_ERR_UNSUPPORTED, "dgc_disc2d_create: order must be in [1, " +
                                                  std::to_string(dgc::kMaxOrder) + "]");
  DGC_REQUIRE(inv_jac && J && face_J && iface_cell && iface_face && nx && ny,
              "dgc_disc2d_create: null geometry array");
  DGC_REQUIRE(num_boundary_nodes >= 0, "dgc_disc2d_create: negative boundary node count");
  DGC_REQUIRE(num_boundary_nodes == 0 || (bx && by), "dgc_disc2d_create: null boundary nodes");
  const int K = num_cells, Nfp = order + 1;
  // The face node tables are fixed by the node ordering; verify the caller's.
  const int* fn[3] = {face0_nodes, face1_nodes, face2_nodes};
  for (int f = 0; f < 3; ++f) {
    if (fn[f] == nullptr) continue;
    for (int d = 0; d < Nfp; ++d) {
      const int Np = (order + 1) * (order + 2) / 2;
      const int want = f == 0 ? d : (f == 1 ? order + d * order - d * (d - 1) / 2 : Np - 1 - d * (d + 3) / 2);
      DGC_REQUIRE(fn[f][d] == want, "dgc_disc2d_create: face node table does not match the node ordering");
    }
  }
  std::vector<double> geo(size_t(dgc::kGeoComps) * K);
  std::vector<int> nbr(size_t(3) * K);
  for (int c = 0; c < K; ++c) {
    if (!(J[c] > 0.0) || !std::isfinite(J[c]))
      return dgc_set_error(DGC_ERR_INVALID, "dgc_disc2d_create: cell " + std::to_string(c) +
                                                " is clockwise or degenerate (J <= 0)");
    geo[size_t(dgc::kRx) * K + c] = inv_jac[4 * size_t(c) + 0];
    geo[size_t(dgc::kRy) * K + c] = inv_jac[4 * size_t(c) + 1];
    geo[size_t(dgc::kSx) * K + c] = inv_jac[4 * size_t(c) + 2];
    geo[size_t(dgc::kSy) * K + c] = inv_jac[4 * size_t(c) + 3];
    for (int e = 0; e < 3; ++e) {
      const size_t h = 3 * size_t(c) + e;
      geo[size_t(dgc::kFs0 + e) * K + c] = face_J[h] / J[c];
      geo[size_t(dgc::kNx0 + e) * K + c] = nx[h];
      geo[size_t(dgc::kNy0 + e) * K + c] = ny[h];
      const int nc = iface_cell[h], nf = iface_face[h];
      if (nc == c) {
        DGC_REQUIRE(nf >= 0 && nf + Nfp <= num_boundary_nodes,
                    "dgc_disc2d_create: boundary offset out of range at cell " + std::to_string(c));
        nbr[size_t(e) * K + c] = (nf << 2) | dgc::kBoundaryCode;
      } else {
        DGC_REQUIRE(nc >= 0 && nc < K && nf >= 0 && nf < 3,
                    "dgc_disc2d_create: neighbour out of range at cell " + std::to_string(c));
        nbr[size_t(e) * K + c] = (nc << 2) | nf;
      }
    }
  }
  DGC_HIP_TRY(hipSetDevice(ctx->device));
  auto* d = new dgc_disc_s;
  d->ctx = ctx;
  d->K = K;
  d->order = order;
  d->Np = (order + 1) * (order + 2) / 2;
  d->Nfp = Nfp;
  d->num_boundary_nodes = num_boundary_nodes;
  d->h_J.assign(J, J + K);
  d->h_face_J.assign(face_J, face_J + 3 * size_t(K));
  d->h_nbr = nbr;
  auto fail = [&](hipError_t e, const char* what) {
    (void)hipFree(d->geo);
    (void)hipFree(d->nbr);
    (void)hipFree(d->J);
    (void)hipFree(d->face_J);
    (void)hipFree(d->bx);
    (void)hipFree(d->by);
    delete d;
    return dgc_set_error(DGC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = hipMalloc(&d->geo, geo.size() * sizeof(double))) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&d->nbr, nbr.size() * sizeof(int))) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&d->J, K * sizeof(double))) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&d->face_J, 3 * size_t(K) * sizeof(double))) != hipSuccess) return fail(e, "hipMalloc");
  const size_t nb = size_t(num_boundary_nodes > 0 ? num_boundary_nodes : 1);
  if ((e = hipMalloc(&d->bx, nb * sizeof(double))) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&d->by, nb * sizeof(double))) != hipSuccess) return fail(e, "hipMalloc");
  std::vector<double> fJ(3 * size_t(K));
  for (int c = 0; c < K; ++c)
    for (int f = 0; f < 3; ++f) fJ[size_t(f) * K + c] = face_J[3 * size_t(c) + f];
  if ((e = hipMemcpy(d->geo, geo.data(), geo.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "hipMemcpy");
  if ((e = hipMemcpy(d->nbr, nbr.data(), nbr.size() * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "hipMemcpy");
  if ((e = hipMemcpy(d->J, J, K * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "hipMemcpy");
  if ((e = hipMemcpy(d->face_J, fJ.data(), fJ.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "hipMemcpy");
  if (num_boundary_nodes > 0) {
    if ((e = hipMemcpy(d->bx, bx, num_boundary_nodes * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
      return fail(e, "hipMemcpy");
    if ((e = hipMemcpy(d->by, by, num_boundary_nodes * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
      return fail(e, "hipMemcpy");
  }
  *out = d;
  return DGC_OK;
}

int dgc_disc2d_destroy(dgc_disc_t disc) {
  if (disc == nullptr) return DGC_OK;
  (void)hipSetDevice(disc->ctx->device);
  (void)hipStreamSynchronize(disc->ctx->stream);
  (void)hipFree(disc->geo);
  (void)hipFree(disc->nbr);
  (void)hipFree(disc->J);
  (void)hipFree(disc->face_J);
  (void)hipFree(disc->bx);
  (void)hipFree(disc->by);
  delete disc;
  return DGC_OK;
}

// ---------------------------------------------------------------------------
static int upload_operators(dgc_op_s* op, const double* Dr, const double* Ds, const double* lift) {
  dgc_disc_s* d = op->disc;
  const int Np = d->Np, NF = 3 * d->Nfp, Npad = dgc::padded_rows(Np);
  std::vector<double> drT(size_t(Np) * Npad, 0.0), dsT(size_t(Np) * Npad, 0.0), lT(size_t(NF) * Npad, 0.0);
  for (int j = 0; j < Np; ++j) {
    for (int k = 0; k < Np; ++k) {
      drT[size_t(k) * Npad + j] = Dr[size_t(j) * Np + k];
      dsT[size_t(k) * Npad + j] = Ds[size_t(j) * Np + k];
    }
    for (int m = 0; m < NF; ++m) lT[size_t(m) * Npad + j] = lift[size_t(j) * NF + m];
  }
  DGC_HIP_TRY(hipMalloc(&op->DrT, drT.size() * sizeof(double)));
  DGC_HIP_TRY(hipMalloc(&op->DsT, dsT.size() * sizeof(double)));
  DGC_HIP_TRY(hipMalloc(&op->LT, lT.size() * sizeof(double)));
  DGC_HIP_TRY(hipMalloc(&op->tclock, 2 * sizeof(double)));
  DGC_HIP_TRY(hipMalloc(&op->flag, sizeof(int)));
  DGC_HIP_TRY(hipMemcpy(op->DrT, drT.data(), drT.size() * sizeof(double), hipMemcpyHostToDevice));
  DGC_HIP_TRY(hipMemcpy(op->DsT, dsT.data(), dsT.size() * sizeof(double), hipMemcpyHostToDevice));
  DGC_HIP_TRY(hipMemcpy(op->LT, lT.data(), lT.size() * sizeof(double), hipMemcpyHostToDevice));
  op->hDr.assign(Dr, Dr + size_t(Np) * Np);
  op->hDs.assign(Ds, Ds + size_t(Np) * Np);
  op->hL.assign(lift, lift + size_t(Np) * NF);
  return DGC_OK;
}

// FP32 operand blocks of the MFMA stage kernel: block a holds rows
// 16 mb .. 16 mb + 15 and columns 4 ks .. 4 ks + 3 of one matrix, lane l
// storing entry (16 mb + l % 16, 4 ks + l / 16) (zero outside the matrix);
// blocks are ordered Dr (mb, ks), Ds (mb, ks), LIFT (mb, ks).
static int build_mfma_operands(dgc_op_s* op) {
  const int Np = op->disc->Np, NF = 3 * op->disc->Nfp;
  const int MB = (Np + 15) / 16, KS = (Np + 3) / 4, KL = (NF + 3) / 4;
  const int NA = dgc::mfma_operand_blocks(Np, op->disc->Nfp);
  std::vector<double> h(size_t(NA) * 64, 0.0);
  auto put = [&](int a, const std::vector<double>& M, int cols, int mb, int ks) {
    for (int l = 0; l < 64; ++l) {
      const int i = 16 * mb + l % 16, k = 4 * ks + l / 16;
      if (i < Np && k < cols) h[size_t(a) * 64 + l] = M[size_t(i) * cols + k];
    }
  };
  for (int mb = 0; mb < MB; ++mb)
    for (int ks = 0; ks < KS; ++ks) {
      put(mb * KS + ks, op->hDr, Np, mb, ks);
      put(MB * KS + mb * KS + ks, op->hDs, Np, mb, ks);
    }
  for (int mb = 0; mb < MB; ++mb)
    for (int ks = 0; ks < KL; ++ks) put(2 * MB * KS + mb * KL + ks, op->hL, NF, mb, ks);
  const std::vector<float> hf(h.begin(), h.end());
  DGC_HIP_TRY(hipMalloc(&op->opD, h.size() * sizeof(double)));
  DGC_HIP_TRY(hipMalloc(&op->opF, hf.size() * sizeof(float)));
  DGC_HIP_TRY(hipMemcpy(op->opD, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  DGC_HIP_TRY(hipMemcpy(op->opF, hf.data(), hf.size() * sizeof(float), hipMemcpyHostToDevice));
  return DGC_OK;
}

static void free_op(dgc_op_s* op) {
  drop_graphs(op);
  (void)hipFree(op->opF);
  (void)hipFree(op->opD);
  (void)hipFree(op->DrT);
  (void)hipFree(op->DsT);
  (void)hipFree(op->LT);
  (void)hipFree(op->tclock);
  (void)hipFree(op->flag);
  delete op;
}

static int create_op(dgc_disc_t disc, int physics, const double* Dr, const double* Ds,
                     const double* lift, double ax, double ay, int bc, dgc_op_t* out) {
  DGC_REQUIRE(disc != nullptr && out != nullptr, "create operator: null argument");
  DGC_REQUIRE(Dr && Ds && lift, "create operator: null operator matrix");
  for (size_t i = 0; i < size_t(disc->Np) * disc->Np; ++i)
    DGC_REQUIRE(std::isfinite(Dr[i]) && std::isfinite(Ds[i]), "create operator: non-finite Dr/Ds");
  DGC_HIP_TRY(hipSetDevice(disc->ctx->device));
  auto* op = new dgc_op_s;
  op->disc = disc;
  op->physics = physics;
  op->F = physics == dgc::kMaxwellTM ? 3 : 1;
  op->ax = ax;
  op->ay = ay;
  op->bc = bc;
  op->wt = dgc::default_waves_per_tile(disc->order, physics, disc->K);
  int rc = upload_operators(op, Dr, Ds, lift);
  if (rc == DGC_OK) rc = build_mfma_operands(op);
  if (rc != DGC_OK) {
    free_op(op);
    return rc;
  }
  *out = op;
  return DGC_OK;
}

int dgc_maxwell_tm_create(dgc_disc_t disc, const double* Dr, const double* Ds, const double* lift,
                          dgc_op_t* out) {
  return create_op(disc, dgc::kMaxwellTM, Dr, Ds, lift, 0.0, 0.0, 0, out);
}

int dgc_advection_create(dgc_disc_t disc, const double* Dr, const double* Ds, const double* lift,
                         double ax, double ay, int bc, dgc_op_t* out) {
  DGC_REQUIRE(bc == dgc::kBcSinSin || bc == dgc::kBcZero, "dgc_advection_create: unknown bc");
  DGC_REQUIRE(std::isfinite(ax) && std::isfinite(ay), "dgc_advection_create: non-finite velocity");
  return create_op(disc, dgc::kAdvection, Dr, Ds, lift, ax, ay, bc, out);
}

int dgc_op_destroy(dgc_op_t op) {
  if (op == nullptr) return DGC_OK;
  (void)hipSetDevice(op->disc->ctx->device);
  (void)hipStreamSynchronize(op->disc->ctx->stream);
  free_op(op);
  return DGC_OK;
}

int dgc_op_num_fields(dgc_op_t op) { return op ? op->F : 0; }

static int op_streaming(const dgc_op_s* op) {
  if (op->cache_policy == DGC_CACHE_AUTO)
    return dgc::rk_working_set_bytes(op->disc->K, op->disc->Np, op->F) > dgc::kLastLevelCacheBytes ? 1 : 0;
  return op->cache_policy == DGC_CACHE_STREAMING ? 1 : 0;
}

static dgc::StageParams base_params(const dgc_op_s* op) {
  dgc::StageParams p{};
  const dgc_disc_s* d = op->disc;
  p.K = d->K;
  p.bc = op->bc;
  p.DrT = op->DrT;
  p.DsT = op->DsT;
  p.LT = op->LT;
  p.geo = d->geo;
  p.nbr = d->nbr;
  p.bx = d->bx;
  p.by = d->by;
  p.ax = op->ax;
  p.ay = op->ay;
  p.tclock = nullptr;
  p.opF = op->precision == 1 ? op->opF : nullptr;
  p.opD = op->opD;
  p.streaming = op_streaming(op);
  return p;
}

static int check_fields(const dgc_op_s* op, const double* const* v, const char* what) {
  if (v == nullptr) return dgc_set_error(DGC_ERR_INVALID, std::string(what) + ": null field array");
  for (int f = 0; f < op->F; ++f)
    if (v[f] == nullptr) return dgc_set_error(DGC_ERR_INVALID, std::string(what) + ": null field pointer");
  return DGC_OK;
}

int dgc_op_rhs(dgc_op_t op, const double* const* in, size_t size, double t, double* const* out) {
  DGC_REQUIRE(op != nullptr, "dgc_op_rhs: null operator");
  DGC_REQUIRE(size == size_t(op->disc->K) * op->disc->Np, "dgc_op_rhs: size != K * Np");
  int rc;
  if ((rc = check_fields(op, in, "dgc_op_rhs")) != DGC_OK) return rc;
  if ((rc = check_fields(op, out, "dgc_op_rhs")) != DGC_OK) return rc;
  dgc::StageParams p = base_params(op);
  p.t = t;
  for (int f = 0; f < op->F; ++f) {
    p.in[f] = in[f];
    p.out[f] = out[f];
  }
  hipError_t e = dgc::launch_stage(op->disc->order, op->physics, 0, op->wt, p, op->disc->ctx->stream);
  if (e != hipSuccess) return dgc_set_error(DGC_ERR_LAUNCH, "failed to launch kernel/device code!");
  return DGC_OK;
}

__global__ void axpy_kernel(double a, const double* __restrict__ x, size_t n,
                            const double* __restrict__ y, double* __restrict__ out) {
  const size_t stride = size_t(blockDim.x) * gridDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = a * x[i] + y[i];
}

// One pair of doubles per thread over a grid of n2/256 workgroups, 16-byte
// nontemporal loads and stores (n2 = number of pairs; arrays 16-byte
// aligned). Measured on MI355X over 3 GiB of traffic: 6.5 TB/s, against
// 4.4-4.8 TB/s for grid-stride loops over 4096-16384 workgroups.
typedef double axpy_v2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void axpy2_kernel(double a, const axpy_v2* __restrict__ x, size_t n2,
                                                    const axpy_v2* __restrict__ y, axpy_v2* __restrict__ out) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n2) return;
  const axpy_v2 xv = __builtin_nontemporal_load(x + i), yv = __builtin_nontemporal_load(y + i);
  __builtin_nontemporal_store(a * xv + yv, out + i);
}

int dgc_axpy(dgc_ctx_t ctx, double a, const double* x, size_t n, const double* y, double* out) {
  DGC_REQUIRE(ctx != nullptr, "dgc_axpy: null context");
  if (n == 0) return DGC_OK;
  DGC_REQUIRE(x && y && out, "dgc_axpy: null pointer");
  const bool aligned = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y) |
                         reinterpret_cast<uintptr_t>(out)) & 15) == 0;
  if (aligned && n >= 2) {
    const size_t n2 = n / 2;
    DGC_REQUIRE((n2 + 255) / 256 < (size_t(1) << 31), "dgc_axpy: vector too long");
    hipLaunchKernelGGL(axpy2_kernel, dim3(unsigned((n2 + 255) / 256)), dim3(256), 0, ctx->stream, a,
                       reinterpret_cast<const axpy_v2*>(x), n2, reinterpret_cast<const axpy_v2*>(y),
                       reinterpret_cast<axpy_v2*>(out));
    if (hipGetLastError() != hipSuccess) return dgc_set_error(DGC_ERR_LAUNCH, "failed to launch kernel/device code!");
    if (n % 2 == 0) return DGC_OK;
    // the odd last element
    x += n - 1;
    y += n - 1;
    out += n - 1;
    n = 1;
  }
  const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(axpy_kernel, dim3(unsigned(blocks)), dim3(256), 0, ctx->stream, a, x, n, y, out);
  if (hipGetLastError() != hipSuccess) return dgc_set_error(DGC_ERR_LAUNCH, "failed to launch kernel/device code!");
  return DGC_OK;
}

// One fused RK4 step. wk mapping: wk0 = stage accumulator (k1 + 2 k2),
// wk1 / wk2 = stage inputs (ping-pong), wk3 = k3. wk4 is unused.
// With `evs` (8 events) each stage launch is bracketed by evs[2s], evs[2s+1].
// The four stage parameter records of one step (stages 1..4 in p[0..3]).
static void step_params(dgc_op_s* op, double* const* u, double t, double dt, double* const* wk, bool use_clock,
                        dgc::StageParams p[4]) {
  const int F = op->F;
  dgc::StageParams b = base_params(op);
  for (int f = 0; f < F; ++f) {
    b.u[f] = u[f];
    b.acc[f] = wk[0 * F + f];
    b.k3[f] = wk[3 * F + f];
  }
  if (use_clock) {
    b.tclock = op->tclock;
    b.dt_clock = dt;
  }
  const double half = 0.5 * dt, sixth = dt / 6.0;
  struct Stage {
    double tfrac;
    double c1;
    int in_buf;   // -1: state, else work index
    int out_buf;  // -1: state
  } stages[4] = {{0.0, half, -1, 1}, {0.5, half, 1, 2}, {0.5, dt, 2, 1}, {1.0, sixth, 1, -1}};
  for (int i = 0; i < 4; ++i) {
    const Stage& s = stages[i];
    p[i] = b;
    p[i].t = (s.tfrac == 0.0) ? t : (s.tfrac == 1.0 ? t + dt : t + 0.5 * dt);
    p[i].tfrac = s.tfrac;
    p[i].c1 = s.c1;
    for (int f = 0; f < F; ++f) {
      p[i].in[f] = s.in_buf < 0 ? u[f] : wk[s.in_buf * F + f];
      p[i].out[f] = s.out_buf < 0 ? u[f] : wk[s.out_buf * F + f];
    }
  }
}

static int enqueue_step(dgc_op_s* op, double* const* u, double t, double dt, double* const* wk,
                        bool use_clock, hipEvent_t* evs = nullptr) {
  dgc::StageParams p[4];
  step_params(op, u, t, dt, wk, use_clock, p);
  hipStream_t st = op->disc->ctx->stream;
  for (int i = 0; i < 4; ++i) {
    if (evs) DGC_HIP_TRY(hipEventRecord(evs[2 * i], st));
    hipError_t e = dgc::launch_stage(op->disc->order, op->physics, i + 1, op->wt, p[i], st);
    if (e != hipSuccess) return dgc_set_error(DGC_ERR_LAUNCH, "failed to launch kernel/device code!");
    if (evs) DGC_HIP_TRY(hipEventRecord(evs[2 * i + 1], st));
  }
  return DGC_OK;
}

__global__ void tick_kernel(double* tclock, double dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) tclock[0] = tclock[0] + dt;
}

static int check_step_args(dgc_op_s* op, double* const* inout, size_t size, double* const* wk,
                           const char* what) {
  if (op == nullptr) return dgc_set_error(DGC_ERR_INVALID, std::string(what) + ": null operator");
  if (size != size_t(op->disc->K) * op->disc->Np)
    return dgc_set_error(DGC_ERR_INVALID, std::string(what) + ": size != K * Np");
  int rc;
  if ((rc = check_fields(op, inout, what)) != DGC_OK) return rc;
  if (wk == nullptr) return dgc_set_error(DGC_ERR_INVALID, std::string(what) + ": null work array");
  for (int i = 0; i < 4 * op->F; ++i)
    if (wk[i] == nullptr) return dgc_set_error(DGC_ERR_INVALID, std::string(what) + ": null work vector");
  return DGC_OK;
}

struct FieldPtrs {
  const double* f[3];
};

__global__ void flag_clear_kernel(int* flag) {
  if (threadIdx.x == 0 && blockIdx.x == 0) flag[0] = 0;
}

// flag |= 1 if any of the F fields of n values holds a NaN or an infinity
// (one atomic per wave that saw one).
__global__ void nonfinite_scan_kernel(FieldPtrs p, int F, size_t n, int* flag) {
  bool bad = false;
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    for (int f = 0; f < F; ++f) bad |= !isfinite(p.f[f][i]);
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

// The time loops' state check (the reference's has_nan helper,
// examples/advection_2d/main.cpp:53-58, applied once per call): one pass over
// the state after the queued steps; blocks until it is known.
static int check_state_finite(dgc_op_s* op, double* const* u, const char* what) {
  hipStream_t s = op->disc->ctx->stream;
  FieldPtrs fp{};
  for (int f = 0; f < op->F; ++f) fp.f[f] = u[f];
  const size_t n = size_t(op->disc->K) * op->disc->Np;
  const int grid = int(std::min<size_t>((n + 255) / 256, 2048));
  hipLaunchKernelGGL(flag_clear_kernel, dim3(1), dim3(64), 0, s, op->flag);
  hipLaunchKernelGGL(nonfinite_scan_kernel, dim3(grid), dim3(256), 0, s, fp, op->F, n, op->flag);
  DGC_HIP_TRY(hipGetLastError());
  int h = 0;
  DGC_HIP_TRY(hipMemcpyAsync(&h, op->flag, sizeof(int), hipMemcpyDeviceToHost, s));
  DGC_HIP_TRY(hipStreamSynchronize(s));
  if (h != 0) return dgc_set_error(DGC_ERR_NONFINITE, std::string(what) + ": the state holds NaN or Inf values");
  return DGC_OK;
}

int dgc_rk4_step(dgc_op_t op, double* const* inout, size_t size, double t, double dt,
                 double* const* wk) {
  int rc = check_step_args(op, inout, size, wk, "dgc_rk4_step");
  if (rc != DGC_OK) return rc;
  return enqueue_step(op, inout, t, dt, wk, false);
}

// dgc_rk4_step with the work vectors in the device's static work area: no
// allocation, no synchronisation. Every stage writes each work vector it
// reads later in the step, so nothing needs clearing between calls.
int dgc_rk4_step_noalloc(dgc_op_t op, double* const* inout, size_t size, double t, double dt) {
  DGC_REQUIRE(op != nullptr, "dgc_rk4_step_noalloc: null operator");
  DGC_REQUIRE(size == size_t(op->disc->K) * op->disc->Np, "dgc_rk4_step_noalloc: size != K * Np");
  int rc = check_fields(op, inout, "dgc_rk4_step_noalloc");
  if (rc != DGC_OK) return rc;
  const int F = op->F;
  if (size_t(4) * F * size > kScratchBytes / sizeof(double))
    return dgc_set_error(DGC_ERR_INVALID, "dgc_rk4_step_noalloc: 4 * F * K * Np = " +
                                              std::to_string(size_t(4) * F * size) +
                                              " doubles exceed the 64 MiB work area (F * K * Np <= 2097152)");
  double* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = g_scratch_ptr[op->disc->ctx->device];
  }
  if (base == nullptr) return dgc_set_error(DGC_ERR_INVALID, "dgc_rk4_step_noalloc: work area not set up");
  double* wk[12] = {};
  for (int i = 0; i < 4 * F; ++i) wk[i] = base + size_t(i) * size;
  return enqueue_step(op, inout, t, dt, wk, false);
}

// Only the advection operator with the analytic inflow data reads the time.
static bool op_uses_time(const dgc_op_s* op) {
  return op->physics == dgc::kAdvection && op->bc == dgc::kBcSinSin && op->disc->num_boundary_nodes > 0;
}

// Build (or reuse) a graph holding `steps` full steps. An operator that reads
// the time runs on the device clock: each step's kernels read the time the
// previous step's tick left.
static int step_graph(dgc_op_s* op, StepGraph& sg, int steps, double* const* u, double dt,
                      double* const* wk) {
  std::vector<const void*> key;
  for (int f = 0; f < op->F; ++f) key.push_back(u[f]);
  for (int i = 0; i < 4 * op->F; ++i) key.push_back(wk[i]);
  if (sg.exec && key == sg.key && sg.dt == dt) return DGC_OK;
  sg.reset();
  hipStream_t s = op->disc->ctx->stream;
  hipGraph_t g = nullptr;
  DGC_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = DGC_OK;
  const bool clock = op_uses_time(op);
  for (int i = 0; i < steps && rc == DGC_OK; ++i) {
    rc = enqueue_step(op, u, 0.0, dt, wk, clock);
    if (rc == DGC_OK && clock) {
      hipLaunchKernelGGL(tick_kernel, dim3(1), dim3(64), 0, s, op->tclock, dt);
    }
  }
  hipError_t e = hipStreamEndCapture(s, &g);
  if (rc != DGC_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return dgc_set_error(DGC_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&sg.exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    sg.exec = nullptr;
    return dgc_set_error(DGC_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  sg.key = key;
  sg.dt = dt;
  return DGC_OK;
}

// Queues nsteps full steps: nsteps / kStepsPerGraph launches of the chunk
// graph, then the remainder from the one-step graph. The kernels and their
// order are those of nsteps one-step launches.
static hipError_t launch_steps(dgc_op_s* op, int nsteps) {
  hipStream_t s = op->disc->ctx->stream;
  hipError_t err = hipSuccess;
  for (int i = 0; i < nsteps / kStepsPerGraph && err == hipSuccess; ++i) err = hipGraphLaunch(op->chunk.exec, s);
  for (int i = 0; i < nsteps % kStepsPerGraph && err == hipSuccess; ++i) err = hipGraphLaunch(op->graph.exec, s);
  return err;
}

// Captures the graphs launch_steps(op, nsteps) replays.
static int step_graphs_for(dgc_op_s* op, int nsteps, double* const* u, double dt, double* const* wk) {
  if (nsteps >= kStepsPerGraph) {
    int rc = step_graph(op, op->chunk, kStepsPerGraph, u, dt, wk);
    if (rc != DGC_OK) return rc;
  }
  if (nsteps % kStepsPerGraph != 0) return step_graph(op, op->graph, 1, u, dt, wk);
  return DGC_OK;
}

static int run_full_steps(dgc_op_s* op, double* const* u, double t0, double dt, int nsteps,
                          double* const* wk) {
  if (nsteps <= 0) return DGC_OK;
  int rc = step_graphs_for(op, nsteps, u, dt, wk);
  if (rc != DGC_OK) return rc;
  if (op_uses_time(op)) {
    hipStream_t s = op->disc->ctx->stream;
    DGC_HIP_TRY(hipMemcpyAsync(op->tclock, &t0, sizeof(double), hipMemcpyHostToDevice, s));
    // the host copy must stay alive until the copy has run
    DGC_HIP_TRY(hipStreamSynchronize(s));
  }
  DGC_HIP_TRY(launch_steps(op, nsteps));
  return DGC_OK;
}

int dgc_rk4_steps(dgc_op_t op, double* const* inout, size_t size, double t0, double dt, int nsteps,
                  double* const* wk) {
  int rc = check_step_args(op, inout, size, wk, "dgc_rk4_steps");
  if (rc != DGC_OK) return rc;
  DGC_REQUIRE(nsteps >= 0, "dgc_rk4_steps: negative step count");
  DGC_REQUIRE(std::isfinite(dt) && dt > 0.0, "dgc_rk4_steps: dt must be positive");
  rc = run_full_steps(op, inout, t0, dt, nsteps, wk);
  if (rc != DGC_OK) return rc;
  return check_state_finite(op, inout, "dgc_rk4_steps");
}

int dgc_rk4_steps_timed(dgc_op_t op, double* const* inout, size_t size, double t0, double dt,
                        int nsteps, double* const* wk, float* ms) {
  int rc = check_step_args(op, inout, size, wk, "dgc_rk4_steps_timed");
  if (rc != DGC_OK) return rc;
  DGC_REQUIRE(nsteps > 0 && ms != nullptr, "dgc_rk4_steps_timed: need nsteps > 0 and ms");
  DGC_REQUIRE(std::isfinite(dt) && dt > 0.0, "dgc_rk4_steps_timed: dt must be positive");
  rc = step_graphs_for(op, nsteps, inout, dt, wk);
  if (rc != DGC_OK) return rc;
  hipStream_t s = op->disc->ctx->stream;
  DGC_HIP_TRY(hipMemcpyAsync(op->tclock, &t0, sizeof(double), hipMemcpyHostToDevice, s));
  DGC_HIP_TRY(hipStreamSynchronize(s));
  dgc::EventPair ev;
  DGC_HIP_TRY(ev.create());
  hipError_t err = hipEventRecord(ev.start, s);
  if (err == hipSuccess) err = launch_steps(op, nsteps);
  if (err == hipSuccess) err = hipEventRecord(ev.stop, s);
  if (err == hipSuccess) err = hipEventSynchronize(ev.stop);
  if (err == hipSuccess) err = hipEventElapsedTime(ms, ev.start, ev.stop);
  if (err != hipSuccess) return dgc_set_error(DGC_ERR_HIP, std::string("timed replay: ") + hipGetErrorString(err));
  return check_state_finite(op, inout, "dgc_rk4_steps_timed");
}

int dgc_op_set_waves_per_tile(dgc_op_t op, int wt) {
  DGC_REQUIRE(op != nullptr, "dgc_op_set_waves_per_tile: null operator");
  DGC_REQUIRE(wt == 1 || wt == 2 || wt == 4 || wt == 8 || wt == 16 || wt == dgc::kSplitTiling ||
                  wt == dgc::kMfmaTiling,
              "dgc_op_set_waves_per_tile: wt must be 1, 2, 4, 8, 16, 32 (rows of a tile split over 4 "
              "workgroups) or 0 (FP64 matrix-core variant)");
  op->wt = dgc::supported_waves_per_tile(op->disc->order, op->physics, wt);
  drop_graphs(op);
  return DGC_OK;
}

int dgc_op_waves_per_tile(dgc_op_t op) { return op ? op->wt : 0; }



int dgc_op_set_precision(dgc_op_t op, int precision) {
  DGC_REQUIRE(op != nullptr, "dgc_op_set_precision: null operator");
  DGC_REQUIRE(precision == 0 || precision == 1, "dgc_op_set_precision: precision must be 0 (FP64) or 1 (FP32 MFMA)");
  op->precision = precision;
  drop_graphs(op);
  return DGC_OK;
}

int dgc_op_precision(dgc_op_t op) { return op ? op->precision : -1; }

int dgc_op_set_cache_policy(dgc_op_t op, int policy) {
  DGC_REQUIRE(op != nullptr, "dgc_op_set_cache_policy: null operator");
  DGC_REQUIRE(policy == DGC_CACHE_AUTO || policy == DGC_CACHE_TEMPORAL || policy == DGC_CACHE_STREAMING,
              "dgc_op_set_cache_policy: policy must be DGC_CACHE_AUTO, DGC_CACHE_TEMPORAL or DGC_CACHE_STREAMING");
  op->cache_policy = policy;
  drop_graphs(op);
  return DGC_OK;
}

int dgc_op_cache_streaming(dgc_op_t op) { return op ? op_streaming(op) : -1; }

// Times every supported tiling on the caller's buffers and keeps the one with
// the shortest RK4 step: the four stage kernels in sequence, as a step runs
// them, except that stage 4 updates wk2 (free once stage 3 has read it) in
// place of the state: same reads and writes. Only work vectors wk0..wk3 are
// written, which the next step's stage 1 re-initialises, so the state is left
// untouched. ms_out[i] receives the mean stage-2 kernel time of candidate i
// (timed on its own).
// Stage parameters of the timing entry points: stage s reads and writes what
// it does in one RK4 step, except that stage 4 updates wk2 (free once stage 3
// has read it) in place of the state; with c1 = 0 only work vectors change.
static void timing_params(dgc_op_s* op, double* const* inout, double* const* wk, dgc::StageParams p[5]) {
  const int F = op->F;
  for (int s = 1; s <= 4; ++s) {
    p[s] = base_params(op);
    p[s].c1 = 0.0;
    for (int f = 0; f < F; ++f) {
      p[s].u[f] = s == 4 ? wk[2 * F + f] : inout[f];
      p[s].acc[f] = wk[0 * F + f];
      p[s].k3[f] = wk[3 * F + f];
      p[s].in[f] = s == 1 ? inout[f] : wk[(s == 3 ? 2 : 1) * F + f];
      p[s].out[f] = wk[(s == 2 ? 2 : 1) * F + f];
    }
  }
}

int dgc_op_stage_time(dgc_op_t op, double* const* inout, size_t size, double* const* wk, int stage, int reps,
                      float* ms) {
  int rc = check_step_args(op, inout, size, wk, "dgc_op_stage_time");
  if (rc != DGC_OK) return rc;
  DGC_REQUIRE(stage >= 1 && stage <= 4 && reps > 0 && ms != nullptr, "dgc_op_stage_time: bad stage, reps or ms");
  dgc::StageParams p[5];
  timing_params(op, inout, wk, p);
  hipStream_t st = op->disc->ctx->stream;
  const int order = op->disc->order, phys = op->physics;
  dgc::EventPair ev;
  DGC_HIP_TRY(ev.create());
  hipError_t e = dgc::launch_stage(order, phys, stage, op->wt, p[stage], st);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(ev.start, st);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = dgc::launch_stage(order, phys, stage, op->wt, p[stage], st);
  if (e == hipSuccess) e = hipEventRecord(ev.stop, st);
  if (e == hipSuccess) e = hipEventSynchronize(ev.stop);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, ev.start, ev.stop);
  if (e != hipSuccess) return dgc_set_error(DGC_ERR_LAUNCH, std::string("dgc_op_stage_time: ") + hipGetErrorString(e));
  *ms = t / reps;
  return DGC_OK;
}

int dgc_op_autotune(dgc_op_t op, double* const* inout, size_t size, double* const* wk, int reps,
                    float* ms_out /* [4]: wt = 1, 2, 4, 0; -1 if not built */) {
  int rc = check_step_args(op, inout, size, wk, "dgc_op_autotune");
  if (rc != DGC_OK) return rc;
  DGC_REQUIRE(reps > 0, "dgc_op_autotune: reps must be positive");
  dgc::StageParams p[5];
  timing_params(op, inout, wk, p);
  hipStream_t st = op->disc->ctx->stream;
  // 8 and 16 waves per tile (one tile per 512- or 1024-thread workgroup) and
  // the split tiling: timed like the others, not reported in ms_out
  const int ncand = 7;
  const int cand[7] = {1, 2, 4, dgc::kMfmaTiling, 8, 16, dgc::kSplitTiling};
  float best = 0.f;
  int best_wt = op->wt;
  hipEvent_t e0, e1, e2;
  DGC_HIP_TRY(hipEventCreate(&e0));
  DGC_HIP_TRY(hipEventCreate(&e1));
  DGC_HIP_TRY(hipEventCreate(&e2));
  const int order = op->disc->order, phys = op->physics;
  for (int i = 0; i < ncand; ++i) {
    const int wt = cand[i];
    if (ms_out && i < 4) ms_out[i] = -1.f;
    if (dgc::supported_waves_per_tile(order, phys, wt) != wt) continue;
    if (op->precision == 1 && wt != 4) continue;  // the matrix-core kernel has one tiling
    // warm-up step, then reps launches of stage 2, then reps steps
    hipError_t e = hipSuccess;
    for (int s = 1; s <= 4 && e == hipSuccess; ++s) e = dgc::launch_stage(order, phys, s, wt, p[s], st);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = dgc::launch_stage(order, phys, 2, wt, p[2], st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    for (int r = 0; r < reps && e == hipSuccess; ++r)
      for (int s = 1; s <= 4 && e == hipSuccess; ++s) e = dgc::launch_stage(order, phys, s, wt, p[s], st);
    if (e == hipSuccess) e = hipEventRecord(e2, st);
    if (e == hipSuccess) e = hipEventSynchronize(e2);
    float ms2 = 0.f, ms_step = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms2, e0, e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms_step, e1, e2);
    if (e != hipSuccess) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      (void)hipEventDestroy(e2);
      return dgc_set_error(DGC_ERR_LAUNCH, std::string("dgc_op_autotune: ") + hipGetErrorString(e));
    }
    ms2 /= reps;
    if (ms_out && i < 4) ms_out[i] = ms2;
    const float step = ms_step / reps;
    if (best == 0.f || step < best) {
      best = step;
      best_wt = wt;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  return dgc_op_set_waves_per_tile(op, best_wt);
}

int dgc_rk4_profile_stages(dgc_op_t op, double* const* inout, size_t size, double t0, double dt,
                           int nsteps, double* const* wk, float* stage_ms) {
  int rc = check_step_args(op, inout, size, wk, "dgc_rk4_profile_stages");
  if (rc != DGC_OK) return rc;
  DGC_REQUIRE(nsteps > 0 && stage_ms != nullptr, "dgc_rk4_profile_stages: need nsteps > 0 and stage_ms");
  hipEvent_t ev[8];
  for (int i = 0; i < 8; ++i) DGC_HIP_TRY(hipEventCreate(&ev[i]));
  double acc[4] = {0, 0, 0, 0};
  double t = t0;
  for (int n = 0; n < nsteps && rc == DGC_OK; ++n) {
    rc = enqueue_step(op, inout, t, dt, wk, false, ev);
    if (rc != DGC_OK) break;
    hipError_t e = hipEventSynchronize(ev[7]);
    for (int sidx = 0; sidx < 4 && e == hipSuccess; ++sidx) {
      float ms = 0.f;
      e = hipEventElapsedTime(&ms, ev[2 * sidx], ev[2 * sidx + 1]);
      acc[sidx] += ms;
    }
    if (e != hipSuccess) rc = dgc_set_error(DGC_ERR_HIP, std::string("stage events: ") + hipGetErrorString(e));
    t += dt;
  }
  for (int i = 0; i < 8; ++i) (void)hipEventDestroy(ev[i]);
  if (rc != DGC_OK) return rc;
  for (int sidx = 0; sidx < 4; ++sidx) stage_ms[sidx] = float(acc[sidx] / nsteps);
  return DGC_OK;
}

int dgc_rk4_run(dgc_op_t op, double* const* inout, size_t size, double t0, double t_end, double dt,
                double* const* wk, double* t_out, int* steps_out) {
  int rc = check_step_args(op, inout, size, wk, "dgc_rk4_run");
  if (rc != DGC_OK) return rc;
  DGC_REQUIRE(std::isfinite(dt) && dt > 0.0, "dgc_rk4_run: dt must be positive");
  DGC_REQUIRE(std::isfinite(t0) && std::isfinite(t_end), "dgc_rk4_run: non-finite time");
  // Replay the host loop `while (t < T) { if (t + dt > T) dt = T - t; step; t += dt; }`
  // to count full steps; the device clock accumulates the same sums.
  double t = t0;
  int nfull = 0;
  while (t < t_end && !(t + dt > t_end)) {
    t += dt;
    ++nfull;
  }
  rc = run_full_steps(op, inout, t0, dt, nfull, wk);
  if (rc != DGC_OK) return rc;
  int steps = nfull;
  if (t < t_end) {
    const double last = t_end - t;
    rc = enqueue_step(op, inout, t, last, wk, false);
    if (rc != DGC_OK) return rc;
    t += last;
    ++steps;
  }
  if (t_out) *t_out = t;
  if (steps_out) *steps_out = steps;
  return check_state_finite(op, inout, "dgc_rk4_run");
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched ensemble of independent Maxwell TM instances.
struct EnsembleGroup {
  int order = 0, wt = 4, offset = 0, nitems = 0;
  bool streaming = false;  // nontemporal accumulator streams
  hipStream_t stream = nullptr;  // branch of the step graph
};

struct dgc_ensemble_s {
  dgc_ctx_s* ctx = nullptr;
  std::vector<dgc_op_s*> ops;
  std::vector<std::vector<double*>> state, work;  // per instance: F and 4F device vectors
  std::vector<double> dt;
  std::vector<int> wt;
  std::vector<EnsembleGroup> groups;  // one per order present
  dgc::StageParams* d_params = nullptr;  // [instance][stage 1..4]
  int2* d_items = nullptr;
  int nitems = 0;
  int cache_policy = DGC_CACHE_AUTO;
  hipGraphExec_t graph = nullptr;
};

static bool ensemble_streaming(const dgc_ensemble_s* e) {
  if (e->cache_policy != DGC_CACHE_AUTO) return e->cache_policy == DGC_CACHE_STREAMING;
  size_t ws = 0;
  for (const dgc_op_s* op : e->ops) ws += dgc::rk_working_set_bytes(op->disc->K, op->disc->Np, op->F);
  return ws > dgc::kLastLevelCacheBytes;
}

static void ensemble_release(dgc_ensemble_s* e) {
  if (e->graph) (void)hipGraphExecDestroy(e->graph);
  e->graph = nullptr;
  for (auto& g : e->groups)
    if (g.stream) (void)hipStreamDestroy(g.stream);
  e->groups.clear();
  (void)hipFree(e->d_params);
  (void)hipFree(e->d_items);
  e->d_params = nullptr;
  e->d_items = nullptr;
}

// Work lists for the current tilings: per order, the instances by decreasing
// size, each as contiguous (instance, tile group) items; and the stage
// parameters of every instance (the buffers of dgc_rk4_step's stages).
static int ensemble_build(dgc_ensemble_s* e) {
  (void)hipStreamSynchronize(e->ctx->stream);
  ensemble_release(e);
  const int n = int(e->ops.size());
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    const dgc_disc_s *da = e->ops[a]->disc, *db = e->ops[b]->disc;
    return da->order != db->order ? da->order > db->order : da->K > db->K;
  });
  std::vector<dgc::StageParams> params(4 * size_t(n));
  for (int i = 0; i < n; ++i) {
    const double h = e->dt[i];
    const double c1[4] = {0.5 * h, 0.5 * h, h, h / 6.0};
    const int in_buf[4] = {-1, 1, 2, 1}, out_buf[4] = {1, 2, 1, -1};
    for (int s = 0; s < 4; ++s) {
      dgc::StageParams p = base_params(e->ops[i]);
      p.c1 = c1[s];
      p.opF = nullptr;
      for (int f = 0; f < 3; ++f) {
        p.u[f] = e->state[i][f];
        p.acc[f] = e->work[i][0 * 3 + f];
        p.k3[f] = e->work[i][3 * 3 + f];
        p.in[f] = in_buf[s] < 0 ? e->state[i][f] : e->work[i][in_buf[s] * 3 + f];
        p.out[f] = out_buf[s] < 0 ? e->state[i][f] : e->work[i][out_buf[s] * 3 + f];
      }
      params[4 * size_t(i) + s] = p;
    }
  }
  std::vector<int2> items;
  for (int i : idx) {
    const int order = e->ops[i]->disc->order;
    if (e->groups.empty() || e->groups.back().order != order) {
      EnsembleGroup g;
      g.order = order;
      g.streaming = ensemble_streaming(e);
      g.wt = e->wt[i];  // the largest instance of the order sets the group's tiling
      g.offset = int(items.size());
      e->groups.push_back(g);
    }
    EnsembleGroup& g = e->groups.back();
    e->wt[i] = g.wt;
    const int cells = g.wt == dgc::kMfmaTiling ? dgc::kTile : dgc::kTile * dgc::tiles_per_group(g.wt);
    const int ngroups = (e->ops[i]->disc->K + cells - 1) / cells;
    for (int t = 0; t < ngroups; ++t) items.push_back(make_int2(i, t));
    g.nitems = int(items.size()) - g.offset;
  }
  for (auto& g : e->groups) DGC_HIP_TRY(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
  e->nitems = int(items.size());
  DGC_HIP_TRY(hipMalloc(&e->d_params, params.size() * sizeof(params[0])));
  DGC_HIP_TRY(hipMalloc(&e->d_items, items.size() * sizeof(int2)));
  DGC_HIP_TRY(hipMemcpy(e->d_params, params.data(), params.size() * sizeof(params[0]), hipMemcpyHostToDevice));
  DGC_HIP_TRY(hipMemcpy(e->d_items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
  return DGC_OK;
}

static int supported_ensemble_wt(int order, int wt) { return dgc::ensemble_supports(order, wt) ? wt : 4; }

extern "C" {

int dgc_ensemble_create(dgc_ctx_t ctx, int n, const dgc_op_t* ops, double* const* states, double* const* work,
                        const double* dt, dgc_ensemble_t* out) {
  DGC_REQUIRE(ctx && ops && states && work && dt && out && n >= 1, "dgc_ensemble_create: bad argument");
  auto* e = new dgc_ensemble_s;
  e->ctx = ctx;
  for (int i = 0; i < n; ++i) {
    dgc_op_s* op = ops[i];
    bool ok = op && op->physics == dgc::kMaxwellTM && op->disc->ctx == ctx && op->precision == 0 &&
              std::isfinite(dt[i]) && dt[i] > 0.0;
    for (int f = 0; ok && f < 3; ++f) ok = states[3 * i + f] != nullptr;
    for (int w = 0; ok && w < 12; ++w) ok = work[12 * i + w] != nullptr;
    if (!ok) {
      delete e;
      return dgc_set_error(DGC_ERR_INVALID, "dgc_ensemble_create: instance " + std::to_string(i) +
                                                " needs a FP64 Maxwell TM operator on this context, dt > 0, "
                                                "3 state and 12 work vectors");
    }
    e->ops.push_back(op);
    e->state.emplace_back(states + 3 * i, states + 3 * i + 3);
    e->work.emplace_back(work + 12 * i, work + 12 * i + 12);
    e->dt.push_back(dt[i]);
    e->wt.push_back(supported_ensemble_wt(op->disc->order, op->wt));
  }
  int rc = ensemble_build(e);
  if (rc != DGC_OK) {
    ensemble_release(e);
    delete e;
    return rc;
  }
  *out = e;
  return DGC_OK;
}

int dgc_ensemble_destroy(dgc_ensemble_t e) {
  if (!e) return DGC_OK;
  (void)hipStreamSynchronize(e->ctx->stream);
  ensemble_release(e);
  delete e;
  return DGC_OK;
}

int dgc_ensemble_autotune(dgc_ensemble_t e, int reps) {
  DGC_REQUIRE(e && reps > 0, "dgc_ensemble_autotune: bad argument");
  const int n = int(e->ops.size());
  for (int order = 1; order <= dgc::kMaxOrder; ++order) {
    int big = -1;
    for (int i = 0; i < n; ++i)
      if (e->ops[i]->disc->order == order && (big < 0 || e->ops[i]->disc->K > e->ops[big]->disc->K)) big = i;
    if (big < 0) continue;
    dgc_op_s* op = e->ops[big];
    float ms[4];
    int rc = dgc_op_autotune(op, e->state[big].data(), size_t(op->disc->K) * op->disc->Np, e->work[big].data(),
                             reps, ms);
    if (rc != DGC_OK) return rc;
    for (int i = 0; i < n; ++i)
      if (e->ops[i]->disc->order == order) e->wt[i] = supported_ensemble_wt(order, op->wt);
  }
  return ensemble_build(e);
}

int dgc_ensemble_info(dgc_ensemble_t e, int* wt, int* num_items, int* grid, size_t* lds_bytes) {
  DGC_REQUIRE(e, "dgc_ensemble_info: null ensemble");
  if (wt) std::copy(e->wt.begin(), e->wt.end(), wt);
  if (num_items) *num_items = e->nitems;
  int g = 0;
  size_t lds = 0;
  for (const auto& gr : e->groups) {
    g += gr.nitems;
    if (gr.wt != dgc::kMfmaTiling)
      lds = std::max(lds, size_t(dgc::stage_lds_doubles(gr.order, dgc::kMaxwellTM, gr.wt)) * sizeof(double));
  }
  if (grid) *grid = g;
  if (lds_bytes) *lds_bytes = lds;
  return DGC_OK;
}

int dgc_ensemble_set_cache_policy(dgc_ensemble_t e, int policy) {
  DGC_REQUIRE(e != nullptr, "dgc_ensemble_set_cache_policy: null ensemble");
  DGC_REQUIRE(policy == DGC_CACHE_AUTO || policy == DGC_CACHE_TEMPORAL || policy == DGC_CACHE_STREAMING,
              "dgc_ensemble_set_cache_policy: policy must be DGC_CACHE_AUTO, DGC_CACHE_TEMPORAL or DGC_CACHE_STREAMING");
  e->cache_policy = policy;
  return ensemble_build(e);
}

int dgc_ensemble_cache_streaming(dgc_ensemble_t e) { return e ? (ensemble_streaming(e) ? 1 : 0) : -1; }

int dgc_ensemble_steps(dgc_ensemble_t e, int nsteps, float* ms) {
  DGC_REQUIRE(e && nsteps >= 0, "dgc_ensemble_steps: bad argument");
  hipStream_t st = e->ctx->stream;
  if (!e->graph) {
    // one step: fork one branch per order, 4 stage launches each, join
    hipGraph_t g = nullptr;
    std::vector<hipEvent_t> evs(e->groups.size() + 1, nullptr);
    hipError_t le = hipSuccess;
    for (auto& ev : evs)
      if (le == hipSuccess) le = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (le == hipSuccess) le = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    const bool capturing = le == hipSuccess;
    if (le == hipSuccess) le = hipEventRecord(evs[0], st);
    for (size_t k = 0; k < e->groups.size() && le == hipSuccess; ++k) {
      const EnsembleGroup& gr = e->groups[k];
      le = hipStreamWaitEvent(gr.stream, evs[0], 0);
      for (int s = 1; s <= 4 && le == hipSuccess; ++s)
        le = dgc::launch_ensemble_stage(gr.order, gr.wt, gr.streaming, s, e->d_params, e->d_items + gr.offset,
                                        gr.nitems, gr.stream);
      if (le == hipSuccess) le = hipEventRecord(evs[k + 1], gr.stream);
      if (le == hipSuccess) le = hipStreamWaitEvent(st, evs[k + 1], 0);
    }
    hipError_t ce = capturing ? hipStreamEndCapture(st, &g) : hipSuccess;
    for (auto& ev : evs)
      if (ev) (void)hipEventDestroy(ev);
    if (le != hipSuccess || ce != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return dgc_set_error(DGC_ERR_LAUNCH, std::string("dgc_ensemble_steps: capture: ") +
                                               hipGetErrorString(le != hipSuccess ? le : ce));
    }
    ce = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ce != hipSuccess) {
      e->graph = nullptr;
      return dgc_set_error(DGC_ERR_HIP, std::string("dgc_ensemble_steps: instantiate: ") + hipGetErrorString(ce));
    }
  }
  dgc::EventPair ev;
  DGC_HIP_TRY(ev.create());
  DGC_HIP_TRY(hipEventRecord(ev.start, st));
  for (int i = 0; i < nsteps; ++i) DGC_HIP_TRY(hipGraphLaunch(e->graph, st));
  DGC_HIP_TRY(hipEventRecord(ev.stop, st));
  DGC_HIP_TRY(hipEventSynchronize(ev.stop));
  float el = 0.f;
  (void)hipEventElapsedTime(&el, ev.start, ev.stop);
  if (ms) *ms = el;
  return DGC_OK;
}

int dgc_ensemble_steps_sequential(dgc_ensemble_t e, int nsteps, float* ms) {
  DGC_REQUIRE(e && nsteps >= 0, "dgc_ensemble_steps_sequential: bad argument");
  hipStream_t st = e->ctx->stream;
  dgc::EventPair ev;
  DGC_HIP_TRY(ev.create());
  DGC_HIP_TRY(hipEventRecord(ev.start, st));
  for (size_t i = 0; i < e->ops.size(); ++i) {
    int rc = run_full_steps(e->ops[i], e->state[i].data(), 0.0, e->dt[i], nsteps, e->work[i].data());
    if (rc != DGC_OK) return rc;
  }
  DGC_HIP_TRY(hipEventRecord(ev.stop, st));
  DGC_HIP_TRY(hipEventSynchronize(ev.stop));
  float el = 0.f;
  (void)hipEventElapsedTime(&el, ev.start, ev.stop);
  if (ms) *ms = el;
  return DGC_OK;
}

}  // extern "C"
