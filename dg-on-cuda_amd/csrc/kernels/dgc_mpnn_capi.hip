// C ABI of the learned preconditioner (include/dgc_hip.h): network state,
// training data from the IPDG operator itself, GPU training with the
// project's forward/backward kernels, and flexible CG with the network and
// a block-Jacobi fallback.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../../include/dgc_hip.h"
#include "../host/mesh.hpp"
#include "dg_internal.hpp"
#include "dg_ipdg.hpp"
#include "dg_mfma.hpp"
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
#define DGC_LAUNCH(call)                                                                                   \
  do {                                                                                                     \
    hipError_t err_ = (call);                                                                              \
    if (err_ != hipSuccess) return dgc_set_error(DGC_ERR_LAUNCH, std::string(#call) + ": " + hipGetErrorString(err_)); \
  } while (0)

struct dgc_mpnn_s {
  dgc_ipdg_s* s = nullptr;
  dgc::MpnnLayout L;
  int K = 0, N = 0, Np = 0;
  float *w = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
  uint16_t* wbf = nullptr;
  uint64_t* wpk[3] = {nullptr, nullptr, nullptr};
  float* partial = nullptr;
  int pstride = 0, nblk = 0, out_grid = 0, cg_grid = 0;
  int prec_grid = 0;  // workgroups (partial pairs) of the preconditioner output
  float* gnorm = nullptr;
  double* mscal = nullptr;
  double* skip_scal = nullptr;  // CG scalars with the done flag set (z = P r only)
  double* tmp_scal = nullptr;   // scratch scalars for reductions outside a solve
  uint16_t* X = nullptr;
  uint16_t* H[3] = {nullptr, nullptr, nullptr};
  float *dh = nullptr, *dh2 = nullptr, *u = nullptr, *vv = nullptr, *dy = nullptr, *y = nullptr;
  float* tgt = nullptr;
  double* loss_partial = nullptr;
  double* partial_out = nullptr;
  double* zfull = nullptr;
  double* dof_x = nullptr;
  double* dof_y = nullptr;
  double* e = nullptr;
  double* amp = nullptr;
  int nsamples = 0;
  double* ds_r = nullptr;         // training residuals r = S e, [nsamples][Np][K]
  std::vector<double> ds_er;      // e . r per sample
  double* wbuf = nullptr;         // S z
  double* dotpart = nullptr;      // [256] + 1 result slot
  int adam_step = 0;
  int base = 1;  // 1: block-Jacobi, 2: two-level (block-Jacobi + coarse correction)
  int op_gen = 0;  // the solver's operator generation the network was built for
  double ct = 1.0;
  hipGraphExec_t graph = nullptr;
  double* graph_hist = nullptr;
  int graph_gen = -1;
  std::vector<void*> owned;
};

namespace {

template <typename T>
int dalloc(dgc_mpnn_s* m, T** p, size_t n) {
  DGC_HIP_TRY(hipMalloc(p, std::max<size_t>(n, 1) * sizeof(T)));
  m->owned.push_back(*p);
  return DGC_OK;
}

dgc::MpnnNet net_of(const dgc_mpnn_s* m) {
  dgc::MpnnNet n;
  n.L = m->L;
  n.K = m->K;
  n.nbr = m->s->disc->nbr;
  n.wbf = m->wbf;
  n.wf = m->w;
  for (int l = 0; l < 3; ++l) n.wpk[l] = m->wpk[l];
  return n;
}

hipStream_t stream_of(const dgc_mpnn_s* m) { return m->s->disc->ctx->stream; }

// Normal deviates from splitmix64 (Box-Muller), deterministic per seed.
struct Gauss {
  dgc_host::SplitMix64 rng;
  explicit Gauss(uint64_t seed) : rng(seed) {}
  double operator()() {
    double u1 = rng.uniform(), u2 = rng.uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

int upload_weights(dgc_mpnn_s* m, const std::vector<float>& w) {
  std::vector<uint16_t> wb(w.size());
  for (size_t i = 0; i < w.size(); ++i) wb[i] = dgc::f2bf(w[i]);
  hipStream_t st = stream_of(m);
  DGC_HIP_TRY(hipMemcpyAsync(m->w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, st));
  DGC_HIP_TRY(hipMemcpyAsync(m->wbf, wb.data(), wb.size() * sizeof(uint16_t), hipMemcpyHostToDevice, st));
  DGC_HIP_TRY(hipMemsetAsync(m->m, 0, w.size() * sizeof(float), st));
  DGC_HIP_TRY(hipMemsetAsync(m->v, 0, w.size() * sizeof(float), st));
  DGC_LAUNCH(dgc::mpnn_pack(m->L, m->wbf, m->wpk, st));
  DGC_HIP_TRY(hipStreamSynchronize(st));
  m->adam_step = 0;
  return DGC_OK;
}

int forward(dgc_mpnn_s* m, const uint16_t* X, const double* cgscal = nullptr) {
  const dgc::MpnnNet net = net_of(m);
  hipStream_t st = stream_of(m);
  DGC_LAUNCH(dgc::mpnn_dense_fwd(net, 0, X, m->H[0], st, cgscal));
  DGC_LAUNCH(dgc::mpnn_dense_fwd(net, 1, m->H[0], m->H[1], st, cgscal));
  DGC_LAUNCH(dgc::mpnn_dense_fwd(net, 2, m->H[1], m->H[2], st, cgscal));
  return DGC_OK;
}

int backward_from(dgc_mpnn_s* m, const uint16_t* X);

// forward + mean-square loss + backward for one sample; gradient in m->grad,
// loss partials in m->loss_partial
int loss_grad(dgc_mpnn_s* m, const uint16_t* X, const float* T) {
  int rc = forward(m, X);
  if (rc) return rc;
  const dgc::MpnnNet net = net_of(m);
  hipStream_t st = stream_of(m);
  dgc::MpnnOutArgs a;
  a.h3 = m->H[2];
  a.t = T;
  a.dy = m->dy;
  a.dh3 = m->dh;
  a.loss_partial = m->loss_partial;
  a.inv_n = 1.0 / (double(m->K) * m->Np);
  DGC_LAUNCH(dgc::mpnn_out(m->N, net, dgc::kOutTrain, a, st));
  return backward_from(m, X);
}

// backward pass from m->dy (dL/dy) and m->dh (dL/dh3) to m->grad
int backward_from(dgc_mpnn_s* m, const uint16_t* X) {
  const dgc::MpnnNet net = net_of(m);
  hipStream_t st = stream_of(m);
  const dgc::MpnnLayout& L = m->L;
  DGC_LAUNCH(dgc::mpnn_wgrad(net, m->Np, m->dy, dgc::kHid, m->H[2], false, m->partial, m->pstride, L.oWo, L.obo, st));
  float* dh = m->dh;
  float* dh2 = m->dh2;
  for (int l = 2; l >= 0; --l) {
    const uint16_t* hin = l == 0 ? X : m->H[l - 1];
    const int din = L.din(l);
    DGC_LAUNCH(dgc::mpnn_dense_bwd(net, l, m->H[l], dh, l > 0 ? m->u : nullptr, l > 0 ? m->vv : nullptr, st));
    DGC_LAUNCH(dgc::mpnn_wgrad(net, dgc::kHid, dh, din, hin, false, m->partial, m->pstride, L.oA[l], L.ob[l], st));
    DGC_LAUNCH(dgc::mpnn_wgrad(net, dgc::kHid, dh, din, hin, true, m->partial, m->pstride, L.oB[l], -1, st));
    if (l > 0) {
      DGC_LAUNCH(dgc::mpnn_nbr_accum(net, din, m->u, m->vv, dh2, st));
      std::swap(dh, dh2);
    }
  }
  DGC_LAUNCH(dgc::mpnn_reduce(m->partial, m->nblk, m->pstride, L.total, m->grad, st));
  return DGC_OK;
}

int host_loss(dgc_mpnn_s* m, double* loss) {
  std::vector<double> lp(m->out_grid);
  DGC_HIP_TRY(hipMemcpyAsync(lp.data(), m->loss_partial, lp.size() * sizeof(double), hipMemcpyDeviceToHost,
                             stream_of(m)));
  DGC_HIP_TRY(hipStreamSynchronize(stream_of(m)));
  *loss = std::accumulate(lp.begin(), lp.end(), 0.0) / (double(m->K) * m->Np);
  return DGC_OK;
}

// zb = P_BJ r, partials (z.z, r.r) in L.partial
int bj_apply(dgc_mpnn_s* m, const double* r, double* zb) {
  IpdgLevel<double>& Lv = m->s->f64;
  dgc::CgParams<double> cp{};
  cp.K = m->K;
  cp.x = Lv.x;
  cp.r = const_cast<double*>(r);
  cp.p = Lv.p;
  cp.q = Lv.q;
  cp.z = zb;
  cp.P = Lv.P;
  cp.scal = m->skip_scal;
  cp.partial = Lv.partial;
  cp.zz = 1;
  DGC_LAUNCH(dgc::cg_update_launch<double>(m->N, cp, stream_of(m)));
  return DGC_OK;
}

// The network's base step zb = P_BJ r (+ the coarse correction for the
// two-level base), partials (z.z, r.r) in L.partial.
int base_apply(dgc_mpnn_s* m, const double* r, double* zb) {
  int rc = bj_apply(m, r, zb);
  if (rc == DGC_OK && m->base == 2) rc = coarse_correct(m->s, r, zb, m->s->f64.partial, 1, nullptr);
  return rc;
}

int precondition(dgc_mpnn_s* m, const double* r, const double* zb, const double* q, const double* cgscal);

// Energy-norm step for one pair (r = S e, er = e.r): z = zb + s_z c_t NN,
// w = S z, loss = (z.w - 2 z.r + er) / er; with `grad`, the parameter
// gradient goes to m->grad. *loss (host) is filled if non-null.
int energy_step(dgc_mpnn_s* m, const double* r, double er, bool grad, double* loss) {
  hipStream_t st = stream_of(m);
  IpdgLevel<double>& Lv = m->s->f64;
  const double n = double(m->K) * m->Np;
  int rc = base_apply(m, r, Lv.z);
  if (rc) return rc;
  DGC_LAUNCH(dgc::fcg_scalars(Lv.partial, m->cg_grid, 3, m->tmp_scal, m->mscal, n, nullptr, st));
  if ((rc = precondition(m, r, Lv.z, nullptr, nullptr))) return rc;
  dgc::IpdgParams<double> pr = Lv.params(*m->s);
  pr.u = m->zfull;
  pr.out = m->wbuf;
  DGC_LAUNCH(dgc::ipdg_apply<double>(m->N, pr, dgc::kSpd, st));
  DGC_LAUNCH(dgc::mpnn_energy_grad(m->N, net_of(m), m->zfull, m->wbuf, r, m->mscal, er, m->dy, m->dh,
                                   m->partial_out, st));
  if (grad && (rc = backward_from(m, m->X))) return rc;
  if (loss) {
    std::vector<double> pp(2 * size_t(m->out_grid));
    DGC_HIP_TRY(hipMemcpyAsync(pp.data(), m->partial_out, pp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    DGC_HIP_TRY(hipStreamSynchronize(st));
    double zw = 0.0, zr = 0.0;
    for (int b = 0; b < m->out_grid; ++b) {
      zw += pp[2 * b];
      zr += pp[2 * b + 1];
    }
    *loss = (zw - 2.0 * zr + er) / er;
  }
  return DGC_OK;
}

// z = zb + s_z c_t NN(r, zb); partials (r.z, z.q) in m->partial_out
int precondition(dgc_mpnn_s* m, const double* r, const double* zb, const double* q, const double* cgscal) {
  const dgc::MpnnNet net = net_of(m);
  hipStream_t st = stream_of(m);
  DGC_LAUNCH(dgc::mpnn_features(m->N, net, r, zb, m->s->disc->geo, m->gnorm, m->mscal, m->X, st, cgscal));
  int rc = forward(m, m->X, cgscal);
  if (rc) return rc;
  dgc::MpnnOutArgs a;
  a.h3 = m->H[2];
  a.zb = zb;
  a.z = m->zfull;
  a.r = r;
  a.q = q;
  a.partial = m->partial_out;
  a.dscal = m->mscal;
  DGC_LAUNCH(dgc::mpnn_out(m->N, net, dgc::kOutPrecond, a, st, cgscal));
  return DGC_OK;
}

// Target scale c_t = RMS((e - zb) / s_z) of one training pair (e in m->e,
// r = S e): the typical size of what the network's output has to supply.
int set_target_scale_from(dgc_mpnn_s* m, const double* r) {
  hipStream_t st = stream_of(m);
  const size_t n = size_t(m->K) * m->Np;
  IpdgLevel<double>& Lv = m->s->f64;
  double* zb = Lv.z;
  int rc = base_apply(m, r, zb);
  if (rc) return rc;
  DGC_LAUNCH(dgc::fcg_scalars(Lv.partial, m->cg_grid, 3, m->tmp_scal, m->mscal, double(n), nullptr, st));
  std::vector<double> he(n), hz(n), ms(dgc::kMpnnScalars);
  DGC_HIP_TRY(hipMemcpyAsync(he.data(), m->e, n * sizeof(double), hipMemcpyDeviceToHost, st));
  DGC_HIP_TRY(hipMemcpyAsync(hz.data(), zb, n * sizeof(double), hipMemcpyDeviceToHost, st));
  DGC_HIP_TRY(hipMemcpyAsync(ms.data(), m->mscal, ms.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  DGC_HIP_TRY(hipStreamSynchronize(st));
  double acc = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double d = (he[i] - hz[i]) / ms[dgc::kSz];
    acc += d * d;
  }
  m->ct = std::sqrt(acc / double(n));
  DGC_REQUIRE(std::isfinite(m->ct) && m->ct > 0, "training data: degenerate target scale");
  ms[dgc::kCt] = m->ct;
  DGC_HIP_TRY(hipMemcpyAsync(m->mscal, ms.data(), ms.size() * sizeof(double), hipMemcpyHostToDevice, st));
  DGC_HIP_TRY(hipStreamSynchronize(st));
  return DGC_OK;
}

// The network, its training data and target scale belong to the operator it
// was created for; a later change of the boundary kind or shift makes them stale.
int check_fresh(const dgc_mpnn_s* m, const char* what) {
  if (m->op_gen != m->s->op_gen || m->s->f64.P == nullptr)
    return dgc_set_error(DGC_ERR_INVALID, std::string(what) +
                                              ": the operator changed (boundary kind or shift) since the network "
                                              "was created; create and train a new network");
  return DGC_OK;
}

}  // namespace

extern "C" {

int dgc_mpnn_create(dgc_ipdg_t s, uint64_t seed, int base, dgc_mpnn_t* out) {
  DGC_REQUIRE(s && out, "dgc_mpnn_create: null argument");
  DGC_REQUIRE(base == 1 || base == 2, "dgc_mpnn_create: base must be 1 (block-Jacobi) or 2 (two-level)");
  DGC_REQUIRE(base == 1 || s->coarse.nc > 0, "dgc_mpnn_create: the two-level base needs dgc_ipdg_coarse_space");
  if (!s->f64.P) {
    int rc = dgc_ipdg_block_jacobi(s, nullptr);
    if (rc) return rc;
  }
  auto* m = new dgc_mpnn_s;
  m->s = s;
  m->base = base;
  m->op_gen = s->op_gen;
  m->K = s->disc->K;
  m->N = s->disc->order;
  m->Np = s->disc->Np;
  m->L = dgc::mpnn_layout(m->Np);
  m->nblk = dgc::mpnn_grid(m->K);
  m->out_grid = (m->K + 255) / 256;
  m->prec_grid = (m->K + dgc::kPrecondCells - 1) / dgc::kPrecondCells;
  m->cg_grid = s->grid;
  m->pstride = m->L.total;
  const size_t K = m->K, P = m->L.total;
  const int dmax = std::max(m->L.D0, dgc::kHid);
  int rc = DGC_OK;
  auto fail = [&](int code) {
    for (void* p : m->owned) (void)hipFree(p);
    delete m;
    return code;
  };
  if ((rc = dalloc(m, &m->w, P)) || (rc = dalloc(m, &m->m, P)) || (rc = dalloc(m, &m->v, P)) ||
      (rc = dalloc(m, &m->grad, P)) || (rc = dalloc(m, &m->wbf, P)) || (rc = dalloc(m, &m->partial, m->nblk * P)) ||
      (rc = dalloc(m, &m->gnorm, 4)) || (rc = dalloc(m, &m->mscal, dgc::kMpnnScalars)) ||
      (rc = dalloc(m, &m->skip_scal, dgc::kCgScalars)) || (rc = dalloc(m, &m->tmp_scal, dgc::kCgScalars)) ||
      (rc = dalloc(m, &m->X, K * m->L.D0)) ||
      (rc = dalloc(m, &m->dh, K * dgc::kHid)) || (rc = dalloc(m, &m->dh2, K * dgc::kHid)) ||
      (rc = dalloc(m, &m->u, K * dmax)) || (rc = dalloc(m, &m->vv, K * dmax)) || (rc = dalloc(m, &m->dy, K * m->Np)) ||
      (rc = dalloc(m, &m->y, K * m->Np)) || (rc = dalloc(m, &m->tgt, K * m->Np)) ||
      (rc = dalloc(m, &m->loss_partial, m->out_grid)) || (rc = dalloc(m, &m->partial_out, 2 * std::max(m->out_grid, m->prec_grid))) ||
      (rc = dalloc(m, &m->zfull, K * m->Np)) || (rc = dalloc(m, &m->amp, 32 * 32)) ||
      (rc = dalloc(m, &m->wbuf, K * m->Np)) || (rc = dalloc(m, &m->dotpart, 257)))
    return fail(rc);
  for (int l = 0; l < 3; ++l) {
    if ((rc = dalloc(m, &m->H[l], K * dgc::kHid))) return fail(rc);
    if ((rc = dalloc(m, &m->wpk[l], size_t(m->L.din(l)) / 4 * 2 * 64))) return fail(rc);
  }
  // feature normalisation of the metric terms: their RMS over the cells
  std::vector<double> geo(4 * K);
  if (hipMemcpy(geo.data(), s->disc->geo, geo.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(dgc_set_error(DGC_ERR_HIP, "dgc_mpnn_create: geometry readback failed"));
  float gn[4];
  for (int k = 0; k < 4; ++k) {
    double a = 0.0;
    for (size_t c = 0; c < K; ++c) a += geo[k * K + c] * geo[k * K + c];
    gn[k] = float(std::sqrt(a / K));
    if (!(gn[k] > 0.f)) gn[k] = 1.f;
  }
  double sk[dgc::kCgScalars] = {0, 0, 0, 0, 0, 1, 0, 0, 0, 0};
  double ms[dgc::kMpnnScalars] = {1.0, 1.0, 1.0, 0.0};
  if (hipMemcpy(m->gnorm, gn, sizeof(gn), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->skip_scal, sk, sizeof(sk), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->mscal, ms, sizeof(ms), hipMemcpyHostToDevice) != hipSuccess)
    return fail(dgc_set_error(DGC_ERR_HIP, "dgc_mpnn_create: upload failed"));
  // He initialisation, deterministic
  std::vector<float> w(P, 0.f);
  Gauss g(seed);
  const dgc::MpnnLayout& L = m->L;
  for (int l = 0; l < 3; ++l) {
    const int d = L.din(l);
    const double sd = std::sqrt(2.0 / (2.0 * d));
    for (int i = 0; i < dgc::kHid * d; ++i) w[L.oA[l] + i] = float(sd * g());
    for (int i = 0; i < dgc::kHid * d; ++i) w[L.oB[l] + i] = float(sd * g());
  }
  // the output layer starts at zero: the untrained network is block-Jacobi
  if ((rc = upload_weights(m, w))) return fail(rc);
  *out = m;
  return DGC_OK;
}

int dgc_mpnn_destroy(dgc_mpnn_t m) {
  if (!m) return DGC_OK;
  (void)hipStreamSynchronize(stream_of(m));
  if (m->graph) (void)hipGraphExecDestroy(m->graph);
  for (void* p : m->owned) (void)hipFree(p);
  (void)hipFree(m->ds_r);
  (void)hipFree(m->dof_x);
  (void)hipFree(m->dof_y);
  (void)hipFree(m->e);
  delete m;
  return DGC_OK;
}

int dgc_mpnn_param_count(dgc_mpnn_t m) { return m ? m->L.total : 0; }
int dgc_mpnn_input_width(dgc_mpnn_t m) { return m ? m->L.D0 : 0; }

int dgc_mpnn_get_params(dgc_mpnn_t m, float* w) {
  DGC_REQUIRE(m && w, "dgc_mpnn_get_params: null argument");
  DGC_HIP_TRY(hipMemcpy(w, m->w, m->L.total * sizeof(float), hipMemcpyDeviceToHost));
  return DGC_OK;
}

int dgc_mpnn_set_params(dgc_mpnn_t m, const float* w) {
  DGC_REQUIRE(m && w, "dgc_mpnn_set_params: null argument");
  for (int i = 0; i < m->L.total; ++i) DGC_REQUIRE(std::isfinite(w[i]), "dgc_mpnn_set_params: non-finite weight");
  return upload_weights(m, std::vector<float>(w, w + m->L.total));
}

int dgc_mpnn_eval(dgc_mpnn_t m, const float* X, float* y) {
  DGC_REQUIRE(m && X && y, "dgc_mpnn_eval: null argument");
  const size_t K = m->K, D0 = m->L.D0;
  std::vector<uint16_t> xb(K * D0);
  for (size_t i = 0; i < xb.size(); ++i) xb[i] = dgc::f2bf(X[i]);
  hipStream_t st = stream_of(m);
  DGC_HIP_TRY(hipMemcpyAsync(m->X, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, st));
  int rc = forward(m, m->X);
  if (rc) return rc;
  dgc::MpnnOutArgs a;
  a.h3 = m->H[2];
  a.y = m->y;
  DGC_LAUNCH(dgc::mpnn_out(m->N, net_of(m), dgc::kOutY, a, st));
  DGC_HIP_TRY(hipMemcpyAsync(y, m->y, K * m->Np * sizeof(float), hipMemcpyDeviceToHost, st));
  DGC_HIP_TRY(hipStreamSynchronize(st));
  return DGC_OK;
}

int dgc_mpnn_loss_grad(dgc_mpnn_t m, const float* X, const float* T, double* loss, float* grad) {
  DGC_REQUIRE(m && X && T, "dgc_mpnn_loss_grad: null argument");
  const size_t K = m->K, D0 = m->L.D0;
  std::vector<uint16_t> xb(K * D0);
  for (size_t i = 0; i < xb.size(); ++i) xb[i] = dgc::f2bf(X[i]);
  hipStream_t st = stream_of(m);
  DGC_HIP_TRY(hipMemcpyAsync(m->X, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, st));
  DGC_HIP_TRY(hipMemcpyAsync(m->tgt, T, K * m->Np * sizeof(float), hipMemcpyHostToDevice, st));
  int rc = loss_grad(m, m->X, m->tgt);
  if (rc) return rc;
  if (grad) DGC_HIP_TRY(hipMemcpyAsync(grad, m->grad, m->L.total * sizeof(float), hipMemcpyDeviceToHost, st));
  if (loss) {
    rc = host_loss(m, loss);
    if (rc) return rc;
  }
  DGC_HIP_TRY(hipStreamSynchronize(st));
  return DGC_OK;
}

int dgc_mpnn_target_scale(dgc_mpnn_t m, double* ct) {
  DGC_REQUIRE(m && ct, "dgc_mpnn_target_scale: null argument");
  *ct = m->ct;
  return DGC_OK;
}

int dgc_mpnn_set_target_scale(dgc_mpnn_t m, double ct) {
  DGC_REQUIRE(m && std::isfinite(ct) && ct > 0, "dgc_mpnn_set_target_scale: need a finite ct > 0");
  hipStream_t st = stream_of(m);
  m->ct = ct;
  DGC_HIP_TRY(hipMemcpyAsync(m->mscal + dgc::kCt, &m->ct, sizeof(double), hipMemcpyHostToDevice, st));
  DGC_HIP_TRY(hipStreamSynchronize(st));
  return DGC_OK;
}

int dgc_mpnn_energy_loss_grad(dgc_mpnn_t m, const double* r, double er, double* loss, float* grad, double* z) {
  DGC_REQUIRE(m && r, "dgc_mpnn_energy_loss_grad: null argument");
  if (int rc_ = check_fresh(m, "dgc_mpnn_energy_loss_grad")) return rc_;
  DGC_REQUIRE(std::isfinite(er) && er > 0, "dgc_mpnn_energy_loss_grad: need e.r > 0");
  hipStream_t st = stream_of(m);
  int rc = energy_step(m, r, er, grad != nullptr, loss);
  if (rc) return rc;
  if (grad) DGC_HIP_TRY(hipMemcpyAsync(grad, m->grad, m->L.total * sizeof(float), hipMemcpyDeviceToHost, st));
  if (z)
    DGC_HIP_TRY(hipMemcpyAsync(z, m->zfull, size_t(m->K) * m->Np * sizeof(double), hipMemcpyDeviceToDevice, st));
  DGC_HIP_TRY(hipStreamSynchronize(st));
  return DGC_OK;
}

int dgc_mpnn_build_dataset(dgc_mpnn_t m, const double* dof_x, const double* dof_y, int nsamples, uint64_t seed0,
                           int modes) {
  DGC_REQUIRE(m && dof_x && dof_y, "dgc_mpnn_build_dataset: null argument");
  if (int rc_ = check_fresh(m, "dgc_mpnn_build_dataset")) return rc_;
  DGC_REQUIRE(nsamples > 0 && modes >= 1 && modes <= 32, "dgc_mpnn_build_dataset: bad nsamples or modes");
  hipStream_t st = stream_of(m);
  const size_t n = size_t(m->K) * m->Np;
  if (!m->dof_x) {
    DGC_HIP_TRY(hipMalloc(&m->dof_x, n * sizeof(double)));
    DGC_HIP_TRY(hipMalloc(&m->dof_y, n * sizeof(double)));
    DGC_HIP_TRY(hipMalloc(&m->e, n * sizeof(double)));
  }
  DGC_HIP_TRY(hipMemcpy(m->dof_x, dof_x, n * sizeof(double), hipMemcpyHostToDevice));
  DGC_HIP_TRY(hipMemcpy(m->dof_y, dof_y, n * sizeof(double), hipMemcpyHostToDevice));
  (void)hipFree(m->ds_r);
  m->ds_r = nullptr;
  m->nsamples = 0;
  DGC_HIP_TRY(hipMalloc(&m->ds_r, size_t(nsamples) * n * sizeof(double)));
  m->ds_er.assign(nsamples, 0.0);
  IpdgLevel<double>& Lv = m->s->f64;
  std::vector<double> amp(size_t(modes) * modes);
  for (int smp = 0; smp < nsamples; ++smp) {
    // error field: sine modes of [-1,1]^2 with N(0,1) k^-2 amplitudes
    Gauss g(seed0 + uint64_t(smp));
    for (int a = 1; a <= modes; ++a)
      for (int b = 1; b <= modes; ++b) amp[size_t(a - 1) * modes + b - 1] = g() / double(a * a + b * b);
    DGC_HIP_TRY(hipMemcpyAsync(m->amp, amp.data(), amp.size() * sizeof(double), hipMemcpyHostToDevice, st));
    DGC_LAUNCH(dgc::grf_eval(m->dof_x, m->dof_y, m->amp, modes, n, m->e, st));
    // residual r = S e through the solver's operator; e.r = |e|_S^2
    double* r = m->ds_r + size_t(smp) * n;
    dgc::IpdgParams<double> pr = Lv.params(*m->s);
    pr.u = m->e;
    pr.out = r;
    DGC_LAUNCH(dgc::ipdg_apply<double>(m->N, pr, dgc::kSpd, st));
    DGC_LAUNCH(dgc::dot_device(m->e, r, n, m->dotpart, m->dotpart + 256, st));
    DGC_HIP_TRY(hipMemcpyAsync(&m->ds_er[smp], m->dotpart + 256, sizeof(double), hipMemcpyDeviceToHost, st));
    if (smp == 0) {
      int rc = set_target_scale_from(m, r);
      if (rc) return rc;
    }
  }
  DGC_HIP_TRY(hipStreamSynchronize(st));
  for (int smp = 0; smp < nsamples; ++smp)
    DGC_REQUIRE(std::isfinite(m->ds_er[smp]) && m->ds_er[smp] > 0, "dgc_mpnn_build_dataset: e.r is not positive");
  m->nsamples = nsamples;
  return DGC_OK;
}

// Training pairs from the solver's own iterations: for right-hand side i
// (b = M f, f the Gaussian random field of seed seed0 + i), PCG with the
// network's base preconditioner runs to 1e-11, recording its iterate x_j
// after every 16 iterations (x_0 = 0); each recorded iterate gives the pair
// e = x* - x_j, r = S e (the residual the preconditioner meets at that point
// of a solve). *npairs receives the number of pairs (up to nrhs * snapshots;
// a solve that converges early contributes fewer).
int dgc_mpnn_build_dataset_solver(dgc_mpnn_t m, const double* dof_x, const double* dof_y, int nrhs,
                                  uint64_t seed0, int modes, int snapshots, int* npairs) {
  DGC_REQUIRE(m && dof_x && dof_y, "dgc_mpnn_build_dataset_solver: null argument");
  if (int rc_ = check_fresh(m, "dgc_mpnn_build_dataset_solver")) return rc_;
  DGC_REQUIRE(nrhs > 0 && snapshots > 0 && modes >= 1 && modes <= 32,
              "dgc_mpnn_build_dataset_solver: bad nrhs, snapshots or modes");
  dgc_ipdg_s* s = m->s;
  DGC_REQUIRE(m->base == 1 || s->coarse.nc > 0, "dgc_mpnn_build_dataset_solver: the coarse space was released");
  hipStream_t st = stream_of(m);
  const size_t n = size_t(m->K) * m->Np;
  if (!m->dof_x) {
    DGC_HIP_TRY(hipMalloc(&m->dof_x, n * sizeof(double)));
    DGC_HIP_TRY(hipMalloc(&m->dof_y, n * sizeof(double)));
    DGC_HIP_TRY(hipMalloc(&m->e, n * sizeof(double)));
  }
  DGC_HIP_TRY(hipMemcpy(m->dof_x, dof_x, n * sizeof(double), hipMemcpyHostToDevice));
  DGC_HIP_TRY(hipMemcpy(m->dof_y, dof_y, n * sizeof(double), hipMemcpyHostToDevice));
  (void)hipFree(m->ds_r);
  m->ds_r = nullptr;
  m->nsamples = 0;
  const size_t cap = size_t(nrhs) * snapshots;
  DGC_HIP_TRY(hipMalloc(&m->ds_r, cap * n * sizeof(double)));
  m->ds_er.assign(cap, 0.0);
  double *snap = nullptr, *b = nullptr, *xs = nullptr;
  struct Buffers {
    double** p[3];
    ~Buffers() {
      for (double** q : p) (void)hipFree(*q);
    }
  } bufs{{&snap, &b, &xs}};
  DGC_HIP_TRY(hipMalloc(&snap, size_t(snapshots) * n * sizeof(double)));
  DGC_HIP_TRY(hipMalloc(&b, n * sizeof(double)));
  DGC_HIP_TRY(hipMalloc(&xs, n * sizeof(double)));
  IpdgLevel<double>& Lv = s->f64;
  std::vector<double> amp(size_t(modes) * modes);
  size_t count = 0;
  for (int i = 0; i < nrhs; ++i) {
    Gauss g(seed0 + uint64_t(i));
    for (int a = 1; a <= modes; ++a)
      for (int c = 1; c <= modes; ++c) amp[size_t(a - 1) * modes + c - 1] = g() / double(a * a + c * c);
    DGC_HIP_TRY(hipMemcpyAsync(m->amp, amp.data(), amp.size() * sizeof(double), hipMemcpyHostToDevice, st));
    DGC_LAUNCH(dgc::grf_eval(m->dof_x, m->dof_y, m->amp, modes, n, m->e, st));
    int rc = dgc_ipdg_mass(s, m->e, b);
    if (rc) return rc;
    DGC_HIP_TRY(hipMemsetAsync(snap, 0, n * sizeof(double), st));  // x_0 = 0
    s->snap = snap + n;
    s->nsnap = snapshots - 1;
    s->snap_taken = 0;
    int it = 0;
    double rel = 0.0;
    float ms = 0.f;
    rc = dgc_ipdg_pcg(s, b, xs, 1e-11, 20000, m->base, &it, &rel, nullptr, &ms);
    const int taken = 1 + s->snap_taken;
    s->snap = nullptr;
    s->nsnap = s->snap_taken = 0;
    if (rc == DGC_ERR_NOT_CONVERGED && rel <= 1e-10) rc = DGC_OK;  // x* to 1e-10 is enough
    if (rc) return rc;
    for (int j = 0; j < taken; ++j) {
      // e = x* - x_j, r = S e, e.r
      double* r = m->ds_r + count * n;
      DGC_LAUNCH(dgc::axpby_device(1.0, xs, -1.0, snap + size_t(j) * n, m->e, n, st));
      dgc::IpdgParams<double> pr = Lv.params(*s);
      pr.u = m->e;
      pr.out = r;
      DGC_LAUNCH(dgc::ipdg_apply<double>(m->N, pr, dgc::kSpd, st));
      DGC_LAUNCH(dgc::dot_device(m->e, r, n, m->dotpart, m->dotpart + 256, st));
      DGC_HIP_TRY(hipMemcpyAsync(&m->ds_er[count], m->dotpart + 256, sizeof(double), hipMemcpyDeviceToHost, st));
      if (count == 0 && (rc = set_target_scale_from(m, r))) return rc;
      ++count;
    }
  }
  DGC_HIP_TRY(hipStreamSynchronize(st));
  for (size_t k = 0; k < count; ++k)
    DGC_REQUIRE(std::isfinite(m->ds_er[k]) && m->ds_er[k] > 0, "dgc_mpnn_build_dataset_solver: e.r is not positive");
  m->ds_er.resize(count);
  m->nsamples = int(count);
  if (npairs) *npairs = int(count);
  return DGC_OK;
}

int dgc_mpnn_train(dgc_mpnn_t m, int first, int count, int epochs, float lr, uint64_t shuffle_seed,
                   double* epoch_loss) {
  DGC_REQUIRE(m, "dgc_mpnn_train: null network");
  if (int rc_ = check_fresh(m, "dgc_mpnn_train")) return rc_;
  DGC_REQUIRE(m->nsamples > 0, "dgc_mpnn_train: build a dataset first");
  DGC_REQUIRE(first >= 0 && count > 0 && first + count <= m->nsamples, "dgc_mpnn_train: sample range");
  DGC_REQUIRE(epochs > 0 && lr > 0.f, "dgc_mpnn_train: bad epochs or learning rate");
  hipStream_t st = stream_of(m);
  const size_t n = size_t(m->K) * m->Np;
  std::vector<int> order(count);
  dgc_host::SplitMix64 rng(shuffle_seed);
  const int total_steps = epochs * count;
  for (int ep = 0; ep < epochs; ++ep) {
    std::iota(order.begin(), order.end(), first);
    for (int i = count - 1; i > 0; --i) std::swap(order[i], order[int(rng.next() % uint64_t(i + 1))]);
    double acc = 0.0;
    for (int k = 0; k < count; ++k) {
      const int smp = order[k];
      double l = 0.0;
      int rc = energy_step(m, m->ds_r + size_t(smp) * n, m->ds_er[smp], true, &l);
      if (rc) return rc;
      if (!std::isfinite(l)) return dgc_set_error(DGC_ERR_NONFINITE, "dgc_mpnn_train: loss is not finite");
      acc += l;
      const int step = ep * count + k;
      // cosine decay to a tenth of the initial rate
      const float lr_t = lr * float(0.1 + 0.9 * 0.5 * (1.0 + std::cos(M_PI * step / double(total_steps))));
      DGC_LAUNCH(dgc::mpnn_adam(m->w, m->m, m->v, m->grad, m->L.total, lr_t, 0.9f, 0.999f, 1e-8f, ++m->adam_step,
                                m->wbf, st));
      DGC_LAUNCH(dgc::mpnn_pack(m->L, m->wbf, m->wpk, st));
    }
    if (epoch_loss) epoch_loss[ep] = acc / count;
  }
  DGC_HIP_TRY(hipStreamSynchronize(st));
  return DGC_OK;
}

int dgc_mpnn_eval_loss(dgc_mpnn_t m, int first, int count, double* loss) {
  DGC_REQUIRE(m && loss, "dgc_mpnn_eval_loss: null argument");
  if (int rc_ = check_fresh(m, "dgc_mpnn_eval_loss")) return rc_;
  DGC_REQUIRE(first >= 0 && count > 0 && first + count <= m->nsamples, "dgc_mpnn_eval_loss: sample range");
  const size_t n = size_t(m->K) * m->Np;
  double acc = 0.0;
  for (int smp = first; smp < first + count; ++smp) {
    double l = 0.0;
    int rc = energy_step(m, m->ds_r + size_t(smp) * n, m->ds_er[smp], false, &l);
    if (rc) return rc;
    acc += l;
  }
  *loss = acc / count;
  return DGC_OK;
}

// Flexible CG (one-term recurrence, p_{k+1} = z_{k+1} - (z_{k+1}.q_k)/(p_k.q_k) p_k)
// with z = zb + s_z c_t NN(r, zb), zb the base step (block-Jacobi or
// two-level). Runs in graph-captured chunks; falls back to PCG with the base
// preconditioner from the current iterate if the residual stalls.
int dgc_ipdg_fcg(dgc_ipdg_t s, dgc_mpnn_t m, const double* b, double* x, double rtol, int maxit, int stall_window,
                 double stall_factor, int* iters, double* relres, int* fell_back, double* history, float* ms) {
  DGC_REQUIRE(s && m && b && x && m->s == s, "dgc_ipdg_fcg: bad argument");
  DGC_REQUIRE(rtol > 0 && maxit > 0, "dgc_ipdg_fcg: need rtol > 0 and maxit > 0");
  DGC_REQUIRE(m->ct > 0, "dgc_ipdg_fcg: the network has no target scale (build a dataset first)");
  DGC_REQUIRE(m->base == 1 || s->coarse.nc > 0, "dgc_ipdg_fcg: the coarse space was released");
  DGC_REQUIRE(m->op_gen == s->op_gen && s->f64.P != nullptr,
              "dgc_ipdg_fcg: the operator changed (boundary kind or shift) since the network was created; "
              "create and train a new network");
  dgc_disc_s* d = s->disc;
  hipStream_t st = d->ctx->stream;
  const size_t n = size_t(d->K) * d->Np;
  IpdgLevel<double>& Lv = s->f64;
  Lv.use_P = true;
  Lv.use_coarse = m->base == 2;
  if (s->hist_len < maxit + 1) {
    (void)hipFree(s->hist);
    s->hist = nullptr;
    DGC_HIP_TRY(hipMalloc(&s->hist, (size_t(maxit) + 1) * sizeof(double)));
    s->hist_len = maxit + 1;
    if (Lv.graph) (void)hipGraphExecDestroy(Lv.graph);
    Lv.graph = nullptr;
    if (m->graph) (void)hipGraphExecDestroy(m->graph);
    m->graph = nullptr;
  }
  // ||b||^2 on the device, as dgc_ipdg_pcg forms it
  double bb = 0.0;
  {
    int rc = dot2(s, b, b, &bb);
    if (rc) return rc;
  }
  DGC_REQUIRE(std::isfinite(bb) && bb > 0, "dgc_ipdg_fcg: zero or non-finite right-hand side");
  const double nd = double(n);

  dgc::EventPair ev;
  DGC_HIP_TRY(ev.create());
  DGC_HIP_TRY(hipEventRecord(ev.start, st));
  // ---- initialisation: r = b, x = 0, z = M(r), p = z
  double h[dgc::kCgScalars] = {0, 0, 0, 0, 0, 0, rtol * rtol * bb, double(maxit), 0, 0};
  DGC_HIP_TRY(hipMemcpyAsync(Lv.scal, h, sizeof(h), hipMemcpyHostToDevice, st));
  DGC_HIP_TRY(hipMemsetAsync(Lv.x, 0, n * sizeof(double), st));
  DGC_HIP_TRY(hipMemsetAsync(Lv.p, 0, n * sizeof(double), st));
  DGC_HIP_TRY(hipMemcpyAsync(Lv.r, b, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  int rc = base_apply(m, Lv.r, Lv.z);
  if (rc) return rc;
  DGC_LAUNCH(dgc::fcg_scalars(Lv.partial, m->cg_grid, 3, Lv.scal, m->mscal, nd, s->hist, st));
  if ((rc = precondition(m, Lv.r, Lv.z, nullptr, Lv.scal))) return rc;
  DGC_LAUNCH(dgc::fcg_scalars(m->partial_out, m->prec_grid, 2, Lv.scal, m->mscal, nd, nullptr, st));
  DGC_LAUNCH(dgc::cg_direction_launch<double>(m->zfull, Lv.p, Lv.scal, n, st));

  // ---- one iteration, captured in chunks
  auto iteration = [&]() -> int {
    dgc::IpdgParams<double> pr = Lv.params(*s);
    pr.u = Lv.p;
    pr.out = Lv.q;
    pr.dotv = Lv.p;
    pr.partial = Lv.partial;
    pr.scal = Lv.scal;
    DGC_LAUNCH(dgc::ipdg_apply<double>(d->order, pr, dgc::kSpd, st));
    DGC_LAUNCH(dgc::cg_scalars_launch<double>(Lv.partial, s->grid, dgc::kAfterApply, Lv.scal, nullptr, st));
    dgc::CgParams<double> cp{};
    cp.K = d->K;
    cp.x = Lv.x;
    cp.r = Lv.r;
    cp.p = Lv.p;
    cp.q = Lv.q;
    cp.z = Lv.z;
    cp.P = Lv.P;
    cp.scal = Lv.scal;
    cp.partial = Lv.partial;
    cp.zz = 1;
    DGC_LAUNCH(dgc::cg_update_launch<double>(d->order, cp, st));
    if (m->base == 2) {
      int r1 = coarse_correct(s, Lv.r, Lv.z, Lv.partial, 1, Lv.scal);
      if (r1) return r1;
    }
    DGC_LAUNCH(dgc::fcg_scalars(Lv.partial, m->cg_grid, 0, Lv.scal, m->mscal, nd, s->hist, st));
    int r2 = precondition(m, Lv.r, Lv.z, Lv.q, Lv.scal);
    if (r2) return r2;
    DGC_LAUNCH(dgc::fcg_scalars(m->partial_out, m->prec_grid, 1, Lv.scal, m->mscal, nd, nullptr, st));
    DGC_LAUNCH(dgc::cg_direction_launch<double>(m->zfull, Lv.p, Lv.scal, n, st));
    return DGC_OK;
  };
  const int chunk = 8;
  if (s->use_graphs && (!m->graph || m->graph_hist != s->hist || m->graph_gen != s->coarse_gen)) {
    if (m->graph) (void)hipGraphExecDestroy(m->graph);
    m->graph = nullptr;
    hipGraph_t g = nullptr;
    DGC_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < chunk && rc == DGC_OK; ++i) rc = iteration();
    hipError_t ce = hipStreamEndCapture(st, &g);
    if (rc != DGC_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (ce != hipSuccess) return dgc_set_error(DGC_ERR_HIP, std::string("fcg capture: ") + hipGetErrorString(ce));
    ce = hipGraphInstantiate(&m->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ce != hipSuccess) {
      m->graph = nullptr;
      return dgc_set_error(DGC_ERR_HIP, std::string("fcg instantiate: ") + hipGetErrorString(ce));
    }
    m->graph_hist = s->hist;
    m->graph_gen = s->coarse_gen;
  }
  // Stall test on the running minimum of the squared residual (CG residuals
  // are not monotone): fall back when the best value of the last
  // stall_window iterations is not below stall_factor^2 times the best value
  // before them.
  std::vector<double> hist_h;
  int it = 0, fb = 0;
  for (;;) {
    DGC_HIP_TRY(hipMemcpyAsync(h, Lv.scal, sizeof(h), hipMemcpyDeviceToHost, st));
    DGC_HIP_TRY(hipStreamSynchronize(st));
    it = int(h[dgc::kIter]);
    if (h[dgc::kDone] != 0.0 || it >= maxit) break;
    if (!(h[dgc::kRr] == h[dgc::kRr])) {
      fb = 1;
      break;
    }
    if (stall_window > 0 && it >= 2 * stall_window) {
      const size_t have = hist_h.size();
      hist_h.resize(it + 1);
      DGC_HIP_TRY(hipMemcpy(hist_h.data() + have, s->hist + have, (it + 1 - have) * sizeof(double),
                            hipMemcpyDeviceToHost));
      const auto split = hist_h.begin() + (it + 1 - stall_window);
      const double best_before = *std::min_element(hist_h.begin(), split);
      const double best_recent = *std::min_element(split, hist_h.end());
      if (best_recent > stall_factor * stall_factor * best_before) {
        fb = 1;
        break;
      }
    }
    if (s->use_graphs) {
      DGC_HIP_TRY(hipGraphLaunch(m->graph, st));
    } else {
      for (int i = 0; i < chunk; ++i)
        if ((rc = iteration())) return rc;
    }
  }
  const bool nan_state = !(h[dgc::kRr] == h[dgc::kRr]);
  if (nan_state) {
    // the network produced non-finite values: restart from zero with the base preconditioner
    DGC_HIP_TRY(hipMemsetAsync(Lv.x, 0, n * sizeof(double), st));
  }
  // true FP64 residual; continue with PCG on the base preconditioner from x
  // after a stall, or restart it while recurrence drift leaves the true
  // residual above rtol
  auto true_rel = [&](double& rel) -> int {
    double rr = 0.0;
    const int r3 = true_residual2(s, b, Lv.x, Lv.q, Lv.r, &rr);
    rel = std::sqrt(rr / bb);
    return r3;
  };
  double rel = 1.0;
  if ((rc = true_rel(rel))) return rc;
  bool more = fb || !(rel <= rtol);
  for (int restart = 0; more && it < maxit && restart < 4; ++restart) {
    if (!fb) fb = 2;  // finished by a PCG restart after recurrence drift
    rc = cg_init_fp64_restart(s, b, rtol * rtol * bb, maxit - it, s->hist + it);
    int it2 = 0;
    if (rc == DGC_OK) rc = cg_run_fp64(s, s->hist + it, maxit - it, &it2);
    if (rc) return rc;
    it += it2;
    if ((rc = true_rel(rel))) return rc;
    more = !(rel <= rtol);
  }
  DGC_HIP_TRY(hipEventRecord(ev.stop, st));
  DGC_HIP_TRY(hipEventSynchronize(ev.stop));
  float el = 0.f;
  (void)hipEventElapsedTime(&el, ev.start, ev.stop);
  DGC_HIP_TRY(hipMemcpyAsync(x, Lv.x, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  if (history) DGC_HIP_TRY(hipMemcpyAsync(history, s->hist, (it + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
  DGC_HIP_TRY(hipStreamSynchronize(st));
  if (iters) *iters = it;
  if (relres) *relres = rel;
  if (fell_back) *fell_back = fb;
  if (ms) *ms = el;
  if (rel > rtol) return dgc_set_error(DGC_ERR_NOT_CONVERGED, "dgc_ipdg_fcg: iteration cap reached");
  return DGC_OK;
}

}  // extern "C"
