// The Stokes dual-splitting step of examples/stokes_2d/stokes_2d.h on the GPU
// (include/dgc_hip.h: dgc_stokes_*), the GPU path the reference leaves as a
// TODO (examples/stokes_2d/k_stokes_2d.cu:96-106, d_stokes_2d.cuh:44-47).
//
// advance_timestep (stokes_2d.h:358-386) from (u^{n-1}, t_prev), (u^n, t) to
// u^{n+1} at t + dt, for the example's analytic problem (a = 2.883356,
// lambda = 1 + a^2; boundary data of stokes_2d.h:100-183):
//   1. u~ = (4 u^n - u^{n-1}) / 3
//   2. pressure: Lap_h p = (3 / 2dt) div u~ + Neumann data h_p (curl curl of the
//      extrapolated velocity plus dg/dt, stokes_2d.h:186-260)
//   3. projection: u~ -= (2 dt / 3) grad p
//   4. viscous: (Lap_h - 3 / 2dt) u = -(3 / 2dt) u~ + Dirichlet data g(t + dt)
// The operators are this library's kernels (grad, div with central fluxes;
// the IPDG Laplacian with Neumann or Dirichlet faces and boundary data). The
// boundary data enter through the whole affine part of the BR1 operator
// (dgc_ipdg_apply_data on a zero field); the reference's
// apply_inhomogeneous_as_rhs leaves out the interior faces of boundary cells,
// which makes its Laplacian with Dirichlet data inconsistent. The reference
// assembles each matrix by unit vectors and solves it with a sparse direct
// solver on the host (stokes_2d.h:388-526); here the
// pressure solve is block-Jacobi PCG on -M Lap_h (singular with Neumann faces:
// the right-hand side is projected onto the complement of the constants), and
// the Helmholtz solves are block-Jacobi PCG on the shifted SPD form
// -M Lap_h + (3 / 2dt) M (dgc_ipdg_set_shift), both with the IPDG solver's
// fused, graph-captured iterations.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../../include/dgc_hip.h"
#include "../host/refelem.hpp"
#include "dg_internal.hpp"
#include "dg_ipdg.hpp"
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
#define DGC_CALL(call)       \
  do {                       \
    int rc_ = (call);        \
    if (rc_ != DGC_OK) return rc_; \
  } while (0)

namespace {

constexpr double kA = 2.883356;
constexpr double kLambda = 1 + kA * kA;

// the example's exact velocity (stokes_2d.h:297-298) at time t
__device__ inline double gx_exact(double x, double y, double t) {
  return sin(x) * (kA * sin(kA * y) - cos(kA) * sinh(y)) * exp(-kLambda * t);
}
__device__ inline double gy_exact(double x, double y, double t) {
  return cos(x) * (cos(kA * y) + cos(kA) * cosh(y)) * exp(-kLambda * t);
}

__global__ void extrapolate_kernel(const double* u0, const double* u1, double* out, size_t n) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) out[i] = (u1[i] * 4. - u0[i]) / 3.;
}

// out = a x + b y
__global__ void axpby_kernel(double a, const double* x, double b, const double* y, double* out, size_t n) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) out[i] = a * x[i] + b * y[i];
}

// Boundary data at the boundary nodes (bx, by):
//   gt = the u~ BC's exterior value at interior 0 (eq. 70 of Fehn et al. 2017,
//        stokes_2d.h:106-114): (2/3) (4 g(t) - g(t_prev))
//   gu = the velocity BC's exterior value at interior 0: 2 g(t_next)
__global__ void boundary_data_kernel(int nb, const double* bx, const double* by, double t, double t_prev,
                                     double t_next, double* gtx, double* gty, double* gux, double* guy) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= nb) return;
  const double x = bx[b], y = by[b];
  const double gUTildeX = 4. * gx_exact(x, y, t) - gx_exact(x, y, t_prev);
  const double gUTildeY = 4. * gy_exact(x, y, t) - gy_exact(x, y, t_prev);
  gtx[b] = 2. * gUTildeX / 3.;
  gty[b] = 2. * gUTildeY / 3.;
  gux[b] = 2. * gx_exact(x, y, t_next);
  guy[b] = 2. * gy_exact(x, y, t_next);
}

// Neumann data of the pressure (stokes_2d.h:186-260), one thread per cell with
// a boundary face: with curl v = Dx v_y - Dy v_x (strong-form Dr, Ds) and
// c = curl u^{n-1} - 2 curl u^n, at boundary node b of face e
//   h_p = dg/dt(t_next).n + (Dy c, -Dx c).n,   hq[b] = 2 h_p
// (the BC's exterior normal derivative at interior value 0).
template <int N>
__global__ __launch_bounds__(256) void pressure_data_kernel(const dgc::IpdgParams<double> p, const double* ux,
                                                            const double* uy, const double* uxp, const double* uyp,
                                                            const double* bx, const double* by, double t_next,
                                                            double* hq) {
  using Tr = dgc::Tri<N>;
  constexpr int Np = Tr::Np, Nfp = Tr::Nfp, Npad = dgc::ipdg_rows(Np);
  const int K = p.K;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= K) return;
  int code[3];
  bool any = false;
  for (int e = 0; e < 3; ++e) {
    code[e] = p.nbr[(size_t)e * K + c];
    any |= (code[e] & 3) == dgc::kBoundaryCode;
  }
  if (!any) return;
  const double rx = p.geo[(size_t)dgc::kRx * K + c], ry = p.geo[(size_t)dgc::kRy * K + c];
  const double sx = p.geo[(size_t)dgc::kSx * K + c], sy = p.geo[(size_t)dgc::kSy * K + c];
  double cc[Np];
  for (int j = 0; j < Np; ++j) {
    double ryn = 0, syn = 0, rxn = 0, sxn = 0, ryp = 0, syp = 0, rxp = 0, sxp = 0;
    for (int k = 0; k < Np; ++k) {
      const double a = p.DrT[k * Npad + j], b = p.DsT[k * Npad + j];
      const size_t i = (size_t)k * K + c;
      ryn += a * uy[i];
      syn += b * uy[i];
      rxn += a * ux[i];
      sxn += b * ux[i];
      ryp += a * uyp[i];
      syp += b * uyp[i];
      rxp += a * uxp[i];
      sxp += b * uxp[i];
    }
    // curl v = (rx Dr + sx Ds) v_y - (ry Dr + sy Ds) v_x
    const double curl_n = (rx * ryn + sx * syn) - (ry * rxn + sy * sxn);
    const double curl_p = (rx * ryp + sx * syp) - (ry * rxp + sy * sxp);
    cc[j] = curl_p - 2. * curl_n;
  }
  for (int e = 0; e < 3; ++e) {
    if ((code[e] & 3) != dgc::kBoundaryCode) continue;
    const int off = code[e] >> 2;
    // outward unit normal of face e from the metric terms
    const double ax = e == 0 ? sx : (e == 1 ? rx + sx : rx);
    const double ay = e == 0 ? sy : (e == 1 ? ry + sy : ry);
    const double l = sqrt(ax * ax + ay * ay);
    const double nx = (e == 1 ? ax : -ax) / l, ny = (e == 1 ? ay : -ay) / l;
    for (int d = 0; d < Nfp; ++d) {
      const int node = Tr::fnode(e, d);
      double dyc = 0, dxc = 0;
      for (int k = 0; k < Np; ++k) {
        const double a = p.DrT[k * Npad + node], b = p.DsT[k * Npad + node];
        dxc += (rx * a + sx * b) * cc[k];
        dyc += (ry * a + sy * b) * cc[k];
      }
      const double curlTerm = dyc * nx - dxc * ny;
      const double x = bx[off + d], y = by[off + d];
      const double dgdtX = kLambda * gx_exact(x, y, t_next);
      const double dgdtY = kLambda * gy_exact(x, y, t_next);
      hq[off + d] = 2. * (dgdtX * nx + dgdtY * ny + curlTerm);
    }
  }
}

hipError_t pressure_data(int order, const dgc::IpdgParams<double>& p, const double* ux, const double* uy,
                         const double* uxp, const double* uyp, const double* bx, const double* by, double t_next,
                         double* hq, hipStream_t s) {
  const dim3 g((p.K + 255) / 256), b(256);
  switch (order) {
    case 1: hipLaunchKernelGGL(pressure_data_kernel<1>, g, b, 0, s, p, ux, uy, uxp, uyp, bx, by, t_next, hq); break;
    case 2: hipLaunchKernelGGL(pressure_data_kernel<2>, g, b, 0, s, p, ux, uy, uxp, uyp, bx, by, t_next, hq); break;
    case 3: hipLaunchKernelGGL(pressure_data_kernel<3>, g, b, 0, s, p, ux, uy, uxp, uyp, bx, by, t_next, hq); break;
    case 4: hipLaunchKernelGGL(pressure_data_kernel<4>, g, b, 0, s, p, ux, uy, uxp, uyp, bx, by, t_next, hq); break;
    case 5: hipLaunchKernelGGL(pressure_data_kernel<5>, g, b, 0, s, p, ux, uy, uxp, uyp, bx, by, t_next, hq); break;
    case 6: hipLaunchKernelGGL(pressure_data_kernel<6>, g, b, 0, s, p, ux, uy, uxp, uyp, bx, by, t_next, hq); break;
    case 7: hipLaunchKernelGGL(pressure_data_kernel<7>, g, b, 0, s, p, ux, uy, uxp, uyp, bx, by, t_next, hq); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

struct dgc_stokes_s {
  dgc_disc_s* disc = nullptr;
  dgc_graddiv_t gd = nullptr;
  dgc_ipdg_t ip_p = nullptr;  // Neumann faces (pressure)
  dgc_ipdg_t ip_u = nullptr;  // Dirichlet faces (velocity)
  std::vector<void*> owned;
  double *ut_x = nullptr, *ut_y = nullptr, *w = nullptr, *b = nullptr, *p = nullptr, *gx = nullptr, *gy = nullptr;
  double *r = nullptr, *d = nullptr, *q = nullptr, *tmp = nullptr, *ones = nullptr;
  double *bx = nullptr, *by = nullptr, *gtx = nullptr, *gty = nullptr, *gux = nullptr, *guy = nullptr, *hq = nullptr;
  int pressure_precond = 1;  // 1: block-Jacobi, 2: two-level (after dgc_stokes_coarse_space)
};

namespace {

size_t n_of(const dgc_stokes_s* s) { return size_t(s->disc->K) * s->disc->Np; }
dim3 grid_of(size_t n) { return dim3(unsigned((n + 255) / 256)); }

template <typename T>
int salloc(dgc_stokes_s* s, T** p, size_t n) {
  if (hipMalloc(reinterpret_cast<void**>(p), (n > 0 ? n : 1) * sizeof(T)) != hipSuccess)
    return dgc_set_error(DGC_ERR_HIP, "dgc_stokes: device allocation failed");
  s->owned.push_back(*p);
  return DGC_OK;
}

// S x = b (S the SPD form of ip) from the initial guess in x, to
// |b - S x| <= rtol |b|: the IPDG solver's PCG (fused, graph-captured
// iterations; precond 1 block-Jacobi, 2 two-level) solves for the correction e
// of r0 = b - S x0 to |r0 - S e| <= (rtol |b| / |r0|) |r0|.
int pcg_from_guess(dgc_stokes_s* s, dgc_ipdg_t ip, int precond, const double* b, double* x, double rtol, int maxit,
                   int* iters) {
  hipStream_t st = s->disc->ctx->stream;
  const size_t n = n_of(s);
  double bb = 0, rr = 0;
  DGC_CALL(dot2(ip, b, b, &bb));
  DGC_CALL(dgc_ipdg_apply_spd(ip, x, s->q));
  hipLaunchKernelGGL(axpby_kernel, grid_of(n), dim3(256), 0, st, 1.0, b, -1.0, (const double*)s->q, s->r, n);
  DGC_HIP_TRY(hipGetLastError());
  DGC_CALL(dot2(ip, s->r, s->r, &rr));
  if (!std::isfinite(rr)) return dgc_set_error(DGC_ERR_NONFINITE, "dgc_stokes: residual is not finite");
  if (bb == 0.0) bb = 1.0;
  int it = 0;
  if (rr > rtol * rtol * bb) {
    double rel = 0;
    float ms = 0;
    DGC_CALL(dgc_ipdg_pcg(ip, s->r, s->d, rtol * std::sqrt(bb / rr), maxit, precond, &it, &rel, nullptr, &ms));
    hipLaunchKernelGGL(axpby_kernel, grid_of(n), dim3(256), 0, st, 1.0, (const double*)s->d, 1.0, (const double*)x,
                       x, n);
    DGC_HIP_TRY(hipGetLastError());
  }
  if (iters) *iters = it;
  return DGC_OK;
}

// (-M Lap_h + sigma M) x = b (Dirichlet faces), block-Jacobi, from the guess in x
int shifted_pcg(dgc_stokes_s* s, double sigma, const double* b, double* x, double rtol, int maxit, int* iters) {
  DGC_CALL(dgc_ipdg_set_shift(s->ip_u, sigma));
  return pcg_from_guess(s, s->ip_u, 1, b, x, rtol, maxit, iters);
}

}  // namespace

extern "C" {

int dgc_stokes_create(dgc_disc_t disc, dgc_stokes_t* out) {
  DGC_REQUIRE(disc != nullptr && out != nullptr, "dgc_stokes_create: null argument");
  const int N = disc->order, Np = disc->Np, NF = 3 * disc->Nfp;
  std::vector<double> Dr(Np * Np), Ds(Np * Np), L(Np * NF), M(Np * Np);
  auto* s = new dgc_stokes_s;
  s->disc = disc;
  auto fail = [&](int rc) {
    dgc_stokes_destroy(s);
    return rc;
  };
  int rc;
  // weak-form operators for grad and div (src/grad_2d.h:66-89), strong-form for the Laplacian
  if ((rc = dgc_ref_operators(N, 0, nullptr, nullptr, nullptr, Dr.data(), Ds.data(), L.data(), nullptr))) return fail(rc);
  if ((rc = dgc_graddiv_create(disc, Dr.data(), Ds.data(), L.data(), &s->gd))) return fail(rc);
  if ((rc = dgc_ref_operators(N, 1, nullptr, nullptr, nullptr, Dr.data(), Ds.data(), L.data(), M.data())))
    return fail(rc);
  if ((rc = dgc_ipdg_create(disc, Dr.data(), Ds.data(), L.data(), M.data(), 0, 1.0, &s->ip_p))) return fail(rc);
  if ((rc = dgc_ipdg_set_boundary(s->ip_p, DGC_BC_NEUMANN))) return fail(rc);
  if ((rc = dgc_ipdg_create(disc, Dr.data(), Ds.data(), L.data(), M.data(), 0, 1.0, &s->ip_u))) return fail(rc);
  const size_t n = size_t(disc->K) * Np, nb = size_t(disc->num_boundary_nodes > 0 ? disc->num_boundary_nodes : 1);
  for (double** v : {&s->ut_x, &s->ut_y, &s->w, &s->b, &s->p, &s->gx, &s->gy, &s->r, &s->d, &s->q, &s->tmp,
                     &s->ones})
    if ((rc = salloc(s, v, n))) return fail(rc);
  for (double** v : {&s->gtx, &s->gty, &s->gux, &s->guy, &s->hq})
    if ((rc = salloc(s, v, nb))) return fail(rc);
  s->bx = disc->bx;
  s->by = disc->by;
  std::vector<double> one(n, 1.0);
  if (hipMemset(s->p, 0, n * sizeof(double)) != hipSuccess ||
      hipMemcpy(s->ones, one.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return fail(dgc_set_error(DGC_ERR_HIP, "dgc_stokes_create: copy failed"));
  *out = s;
  return DGC_OK;
}

int dgc_stokes_coarse_space(dgc_stokes_t s, const double* cell_x, const double* cell_y, int gx, int gy) {
  DGC_REQUIRE(s != nullptr, "dgc_stokes_coarse_space: null solver");
  int nc = 0;
  DGC_CALL(dgc_ipdg_coarse_space(s->ip_p, cell_x, cell_y, gx, gy, &nc));
  s->pressure_precond = 2;
  return DGC_OK;
}

int dgc_stokes_reset(dgc_stokes_t s) {
  DGC_REQUIRE(s != nullptr, "dgc_stokes_reset: null solver");
  DGC_HIP_TRY(hipMemsetAsync(s->p, 0, n_of(s) * sizeof(double), s->disc->ctx->stream));
  return DGC_OK;
}

int dgc_stokes_destroy(dgc_stokes_t s) {
  if (s == nullptr) return DGC_OK;
  (void)hipStreamSynchronize(s->disc->ctx->stream);
  dgc_graddiv_destroy(s->gd);
  dgc_ipdg_destroy(s->ip_p);
  dgc_ipdg_destroy(s->ip_u);
  for (void* p : s->owned) (void)hipFree(p);
  delete s;
  return DGC_OK;
}

int dgc_stokes_advance(dgc_stokes_t s, const double* ux_prev, const double* uy_prev, double t_prev, const double* ux,
                       const double* uy, double t, double dt, double rtol, double* ux_out, double* uy_out,
                       int* iters) {
  DGC_REQUIRE(s && ux_prev && uy_prev && ux && uy && ux_out && uy_out, "dgc_stokes_advance: null argument");
  DGC_REQUIRE(dt > 0 && rtol > 0, "dgc_stokes_advance: need dt > 0 and rtol > 0");
  const dgc_disc_s* d = s->disc;
  hipStream_t st = d->ctx->stream;
  const size_t n = n_of(s);
  const int nb = d->num_boundary_nodes;
  const int maxit = 100000;
  int it_p = 0, it_x = 0, it_y = 0;
  DGC_HIP_TRY(hipSetDevice(d->ctx->device));
  // 1. u~ = (4 u^n - u^{n-1}) / 3; boundary data
  hipLaunchKernelGGL(extrapolate_kernel, grid_of(n), dim3(256), 0, st, ux_prev, ux, s->ut_x, n);
  hipLaunchKernelGGL(extrapolate_kernel, grid_of(n), dim3(256), 0, st, uy_prev, uy, s->ut_y, n);
  if (nb > 0) {
    hipLaunchKernelGGL(boundary_data_kernel, grid_of(size_t(nb)), dim3(256), 0, st, nb, (const double*)s->bx,
                       (const double*)s->by, t, t_prev, t + dt, s->gtx, s->gty, s->gux, s->guy);
    const dgc::IpdgParams<double> pr = s->ip_p->f64.params(*s->ip_p);
    DGC_HIP_TRY(pressure_data(d->order, pr, ux, uy, ux_prev, uy_prev, s->bx, s->by, t + dt, s->hq, st));
  }
  DGC_HIP_TRY(hipGetLastError());
  // 2. pressure: Lap_h p = (3 / 2dt) div u~ + Neumann data  (stokes_2d.h:322-338)
  DGC_CALL(dgc_div2d_apply(s->gd, s->ut_x, s->ut_y, nb ? s->gtx : nullptr, nb ? s->gty : nullptr, -1.0, s->w));
  //    the Neumann data's affine part A(h) of the operator moves to the right: Lap_h p = w - A(h)
  DGC_HIP_TRY(hipMemsetAsync(s->tmp, 0, n * sizeof(double), st));
  DGC_CALL(dgc_ipdg_apply_data(s->ip_p, nullptr, nb ? s->hq : nullptr, s->tmp));
  hipLaunchKernelGGL(axpby_kernel, grid_of(n), dim3(256), 0, st, 1.5 / dt, (const double*)s->w, -1.0,
                     (const double*)s->tmp, s->w, n);
  //    as -M Lap_h p = -M w, projected onto the complement of the constants
  DGC_CALL(dgc_ipdg_mass(s->ip_p, s->w, s->b));
  double sum = 0;
  DGC_CALL(dot2(s->ip_p, s->b, s->ones, &sum));
  hipLaunchKernelGGL(axpby_kernel, grid_of(n), dim3(256), 0, st, -1.0, (const double*)s->b, sum / double(n),
                     (const double*)s->ones, s->b, n);
  //    from the previous step's pressure (zero before the first step)
  DGC_CALL(pcg_from_guess(s, s->ip_p, s->pressure_precond, s->b, s->p, rtol, maxit, &it_p));
  // 3. projection: u~ -= (2 dt / 3) grad p, pressure BC exterior = interior (stokes_2d.h:340-349)
  DGC_CALL(dgc_grad2d_apply(s->gd, s->p, nullptr, 1.0, s->gx, s->gy));
  hipLaunchKernelGGL(axpby_kernel, grid_of(n), dim3(256), 0, st, -2. * dt / 3., (const double*)s->gx, 1.0,
                     (const double*)s->ut_x, s->ut_x, n);
  hipLaunchKernelGGL(axpby_kernel, grid_of(n), dim3(256), 0, st, -2. * dt / 3., (const double*)s->gy, 1.0,
                     (const double*)s->ut_y, s->ut_y, n);
  // 4. viscous: (Lap_h - sigma) u = -sigma u~ + Dirichlet data, sigma = 3 / 2dt (stokes_2d.h:351-368),
  //    as (-M Lap_h + sigma M) u = -M (that right-hand side)
  const double sigma = 1.5 / dt;
  const double* ut[2] = {s->ut_x, s->ut_y};
  const double* gu[2] = {s->gux, s->guy};
  double* uo[2] = {ux_out, uy_out};
  const double* un[2] = {ux, uy};
  const double* up[2] = {ux_prev, uy_prev};
  int* its[2] = {&it_x, &it_y};
  for (int comp = 0; comp < 2; ++comp) {
    // w = -sigma u~ - A(g): the Dirichlet data's whole affine part
    DGC_HIP_TRY(hipMemsetAsync(s->tmp, 0, n * sizeof(double), st));
    DGC_CALL(dgc_ipdg_apply_data(s->ip_u, nb ? gu[comp] : nullptr, nullptr, s->tmp));
    hipLaunchKernelGGL(axpby_kernel, grid_of(n), dim3(256), 0, st, -sigma, ut[comp], -1.0, (const double*)s->tmp,
                       s->w, n);
    DGC_CALL(dgc_ipdg_mass(s->ip_u, s->w, s->b));
    hipLaunchKernelGGL(axpby_kernel, grid_of(n), dim3(256), 0, st, -1.0, (const double*)s->b, 0.0,
                       (const double*)s->b, s->b, n);
    // initial guess: the linear extrapolation 2 u^n - u^{n-1} (closer to u^{n+1}
    // than the projected velocity by a factor of order dt)
    hipLaunchKernelGGL(axpby_kernel, grid_of(n), dim3(256), 0, st, 2.0, un[comp], -1.0, up[comp], uo[comp], n);
    DGC_HIP_TRY(hipGetLastError());
    DGC_CALL(shifted_pcg(s, sigma, s->b, uo[comp], rtol, maxit, its[comp]));
  }
  if (iters) {
    iters[0] = it_p;
    iters[1] = it_x;
    iters[2] = it_y;
  }
  return DGC_OK;
}

}  // extern "C"
