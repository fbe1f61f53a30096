// Device set-up of a discretization from its cells (include/dgc_hip.h,
// dgc_disc2d_create_from_cells): geometric factors, twin half-edges, boundary
// face nodes and the packed connectivity are built on the GPU, for meshes
// where the host set-up would take seconds (K >= 1e6).
//
// The arrays equal the host path's bit for bit (dgc_mesh_geometry followed by
// dgc_disc2d_create): the geometry kernels evaluate the host's expressions in
// the same order with contraction off, and the half-edge sort is stable, so
// twins, boundary offsets and boundary nodes come out in the host's order.
//   geometry:   src/mapping_2d.h:101-158, src/basic_geom_2d.h:110-153
//   twins:      src/mesh/simple_triangular_mesh_2d.h:151-217
//   marshalling src/simple_discretization_2d.h:84-218
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
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

constexpr int kSetupBlock = 256;

// First failing index of each check (INT_MAX: none).
enum SetupError { kErrVertex = 0, kErrEdge = 1, kErrManifold = 2, kErrCell = 3, kErrCount = 4 };

__global__ __launch_bounds__(kSetupBlock) void fill_errors(int* err) {
  if (threadIdx.x < kErrCount) err[threadIdx.x] = INT_MAX;
}

// One thread per cell: metric terms, J, face J, outward normals, and the
// sort keys (min, max vertex id) of its three half-edges h = 3c + e.
__global__ __launch_bounds__(kSetupBlock) void cell_geometry(int K, int Nv, const double* __restrict__ cx,
                                                             const double* __restrict__ cy,
                                                             const int* __restrict__ etov, double* __restrict__ geo,
                                                             double* __restrict__ J, double* __restrict__ fJ,
                                                             double* __restrict__ fJc, uint64_t* __restrict__ key,
                                                             int* __restrict__ val, int* __restrict__ twin,
                                                             int* __restrict__ err) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * kSetupBlock + threadIdx.x;
  if (c >= K) return;
  const size_t Ks = size_t(K);
  const double x1 = cx[3 * size_t(c)], y1 = cy[3 * size_t(c)];
  const double x2 = cx[3 * size_t(c) + 1], y2 = cy[3 * size_t(c) + 1];
  const double x3 = cx[3 * size_t(c) + 2], y3 = cy[3 * size_t(c) + 2];
  // [[x_r, x_s], [y_r, y_s]] and its inverse
  const double xr = 0.5 * (x2 - x1), xs = 0.5 * (x3 - x1);
  const double yr = 0.5 * (y2 - y1), ys = 0.5 * (y3 - y1);
  const double det = xr * ys - xs * yr;
  const double Jc = 0.25 * ((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1));
  if (det == 0.0 || !(Jc > 0.0) || !isfinite(Jc)) atomicMin(err + kErrCell, c);
  geo[size_t(dgc::kRx) * Ks + c] = ys / det;
  geo[size_t(dgc::kRy) * Ks + c] = -xs / det;
  geo[size_t(dgc::kSx) * Ks + c] = -yr / det;
  geo[size_t(dgc::kSy) * Ks + c] = xr / det;
  J[c] = Jc;
  const double vx[3] = {x1, x2, x3}, vy[3] = {y1, y2, y3};
  int ids[3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    ids[v] = etov[3 * size_t(c) + v];
    if (ids[v] < 0 || ids[v] >= Nv) atomicMin(err + kErrVertex, 3 * c + v);
  }
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int a = e, b = (e + 1) % 3, o = (e + 2) % 3;
    double ex = vx[b] - vx[a], ey = vy[b] - vy[a];
    const double fj = 0.5 * sqrt(ex * ex + ey * ey);
    // (ey, -ex) is outward for a counter-clockwise cell; flip if it points at the third vertex
    const double ox = vx[o] - vx[a], oy = vy[o] - vy[a];
    if (ey * ox - ex * oy > 0.0) {
      ex = -ex;
      ey = -ey;
    }
    const double l = sqrt(ex * ex + ey * ey);
    geo[size_t(dgc::kFs0 + e) * Ks + c] = fj / Jc;
    geo[size_t(dgc::kNx0 + e) * Ks + c] = ey / l;
    geo[size_t(dgc::kNy0 + e) * Ks + c] = -ex / l;
    fJ[size_t(e) * Ks + c] = fj;
    fJc[3 * size_t(c) + e] = fj;
    const size_t h = 3 * size_t(c) + e;
    uint64_t p = uint32_t(ids[a]), q = uint32_t(ids[b]);
    if (p == q) atomicMin(err + kErrEdge, int(h));
    if (p > q) {
      const uint64_t t = p;
      p = q;
      q = t;
    }
    key[h] = (p << 32) | q;
    val[h] = int(h);
    twin[h] = -1;
  }
}

// One thread per position of the sorted half-edges: the first of a run of
// equal keys pairs the two twins of an interior edge; a run of one is a
// boundary half-edge (twin stays -1); longer runs are non-manifold edges.
__global__ __launch_bounds__(kSetupBlock) void pair_twins(int n, const uint64_t* __restrict__ skey,
                                                          const int* __restrict__ sval, int* __restrict__ twin,
                                                          int* __restrict__ err) {
  const int i = blockIdx.x * kSetupBlock + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = skey[i];
  if (i > 0 && skey[i - 1] == k) return;
  int run = 1;
  while (i + run < n && run < 3 && skey[i + run] == k) ++run;
  if (run == 2) {
    const int p = sval[i], q = sval[i + 1];
    twin[p] = q;
    twin[q] = p;
  } else if (run > 2) {
    atomicMin(err + kErrManifold, i);
  }
}

__global__ __launch_bounds__(kSetupBlock) void boundary_flags(int n, const int* __restrict__ twin,
                                                              int* __restrict__ flag) {
  const int h = blockIdx.x * kSetupBlock + threadIdx.x;
  if (h < n) flag[h] = twin[h] < 0 ? 1 : 0;
}

// One thread per half-edge: packed connectivity (twin cell << 2 | twin face,
// or boundary node offset << 2 | 3) and, on boundary faces, the physical
// coordinates of the face nodes, numbered in half-edge order.
__global__ __launch_bounds__(kSetupBlock) void pack_faces(int K, int Nfp, const double* __restrict__ cx,
                                                          const double* __restrict__ cy,
                                                          const int* __restrict__ twin,
                                                          const int* __restrict__ pos,
                                                          const double* __restrict__ frs,  // [3][Nfp][r, s]
                                                          int* __restrict__ nbr, double* __restrict__ bx,
                                                          double* __restrict__ by) {
#pragma clang fp contract(off)
  const int h = blockIdx.x * kSetupBlock + threadIdx.x;
  if (h >= 3 * K) return;
  const int c = h / 3, e = h - 3 * (h / 3);
  const int t = twin[h];
  if (t >= 0) {
    nbr[size_t(e) * K + c] = ((t / 3) << 2) | (t % 3);
    return;
  }
  const int offset = pos[h] * Nfp;
  nbr[size_t(e) * K + c] = (offset << 2) | dgc::kBoundaryCode;
  const double x1 = cx[3 * size_t(c)], y1 = cy[3 * size_t(c)];
  const double x2 = cx[3 * size_t(c) + 1], y2 = cy[3 * size_t(c) + 1];
  const double x3 = cx[3 * size_t(c) + 2], y3 = cy[3 * size_t(c) + 2];
  for (int d = 0; d < Nfp; ++d) {
    // rs -> xy with barycentric weights of vertices 1, 2, 3
    const double r = frs[2 * (e * Nfp + d)], s = frs[2 * (e * Nfp + d) + 1];
    const double l1 = 0.5 * (s + 1.0), l2 = -0.5 * (r + s), l3 = 0.5 * (r + 1.0);
    bx[offset + d] = l2 * x1 + l3 * x2 + l1 * x3;
    by[offset + d] = l2 * y1 + l3 * y2 + l1 * y3;
  }
}

// Device buffers of the set-up, released on every exit path.
struct SetupScratch {
  std::vector<void*> blocks;
  template <typename T>
  hipError_t alloc(T** p, size_t count) {
    *p = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), (count > 0 ? count : 1) * sizeof(T));
    if (e == hipSuccess) blocks.push_back(*p);
    return e;
  }
  ~SetupScratch() {
    for (void* b : blocks) (void)hipFree(b);
  }
};

int grid_of(size_t n) { return int((n + kSetupBlock - 1) / kSetupBlock); }

}  // namespace

extern "C" {

int dgc_disc2d_create_from_cells(dgc_ctx_t ctx, int num_cells, int num_vertices, int order, const double* cx,
                                 const double* cy, const int* etov, dgc_disc_t* out) {
  DGC_REQUIRE(ctx != nullptr && out != nullptr, "dgc_disc2d_create_from_cells: null argument");
  DGC_REQUIRE(cx && cy && etov, "dgc_disc2d_create_from_cells: null cell array");
  DGC_REQUIRE(num_cells > 0 && num_cells < (1 << 29), "dgc_disc2d_create_from_cells: num_cells out of range");
  DGC_REQUIRE(num_vertices > 0, "dgc_disc2d_create_from_cells: num_vertices must be positive");
  if (order < 1 || order > dgc::kMaxOrder)
    return dgc_set_error(DGC_ERR_UNSUPPORTED, "dgc_disc2d_create_from_cells: order must be in [1, " +
                                                  std::to_string(dgc::kMaxOrder) + "]");
  const int K = num_cells, Nfp = order + 1, n = 3 * K;
  const size_t Ks = size_t(K), ns = size_t(n);
  hipStream_t st = ctx->stream;
  DGC_HIP_TRY(hipSetDevice(ctx->device));

  // face-node reference coordinates [3][Nfp][r, s]
  std::vector<double> r, s;
  dgc_host::triangle_nodes(order, r, s);
  std::vector<double> frs(size_t(6) * Nfp);
  for (int e = 0; e < 3; ++e)
    for (int d = 0; d < Nfp; ++d) {
      const int node = dgc_host::face_node(order, e, d);
      frs[2 * (e * Nfp + d)] = r[node];
      frs[2 * (e * Nfp + d) + 1] = s[node];
    }

  SetupScratch tmp;
  double *d_cx, *d_cy, *d_frs, *d_fJc;
  int *d_etov, *d_val, *d_sval, *d_twin, *d_flag, *d_pos, *d_err;
  uint64_t *d_key, *d_skey;
  DGC_HIP_TRY(tmp.alloc(&d_cx, ns));
  DGC_HIP_TRY(tmp.alloc(&d_cy, ns));
  DGC_HIP_TRY(tmp.alloc(&d_etov, ns));
  DGC_HIP_TRY(tmp.alloc(&d_frs, frs.size()));
  DGC_HIP_TRY(tmp.alloc(&d_fJc, ns));
  DGC_HIP_TRY(tmp.alloc(&d_key, ns));
  DGC_HIP_TRY(tmp.alloc(&d_skey, ns));
  DGC_HIP_TRY(tmp.alloc(&d_val, ns));
  DGC_HIP_TRY(tmp.alloc(&d_sval, ns));
  DGC_HIP_TRY(tmp.alloc(&d_twin, ns));
  DGC_HIP_TRY(tmp.alloc(&d_flag, ns));
  DGC_HIP_TRY(tmp.alloc(&d_pos, ns));
  DGC_HIP_TRY(tmp.alloc(&d_err, kErrCount));
  DGC_HIP_TRY(hipMemcpyAsync(d_cx, cx, ns * sizeof(double), hipMemcpyHostToDevice, st));
  DGC_HIP_TRY(hipMemcpyAsync(d_cy, cy, ns * sizeof(double), hipMemcpyHostToDevice, st));
  DGC_HIP_TRY(hipMemcpyAsync(d_etov, etov, ns * sizeof(int), hipMemcpyHostToDevice, st));
  DGC_HIP_TRY(hipMemcpyAsync(d_frs, frs.data(), frs.size() * sizeof(double), hipMemcpyHostToDevice, st));

  // the discretization's own arrays (bx, by once the boundary size is known)
  auto* d = new dgc_disc_s;
  d->ctx = ctx;
  d->K = K;
  d->order = order;
  d->Np = (order + 1) * (order + 2) / 2;
  d->Nfp = Nfp;
  auto fail = [&](int code, const std::string& msg) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(d->geo);
    (void)hipFree(d->nbr);
    (void)hipFree(d->J);
    (void)hipFree(d->face_J);
    (void)hipFree(d->bx);
    (void)hipFree(d->by);
    delete d;
    return dgc_set_error(code, msg);
  };
  auto hip_fail = [&](hipError_t e, const char* what) {
    return fail(DGC_ERR_HIP, std::string("dgc_disc2d_create_from_cells: ") + what + ": " + hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = hipMalloc(&d->geo, size_t(dgc::kGeoComps) * Ks * sizeof(double))) != hipSuccess)
    return hip_fail(e, "hipMalloc");
  if ((e = hipMalloc(&d->nbr, ns * sizeof(int))) != hipSuccess) return hip_fail(e, "hipMalloc");
  if ((e = hipMalloc(&d->J, Ks * sizeof(double))) != hipSuccess) return hip_fail(e, "hipMalloc");
  if ((e = hipMalloc(&d->face_J, ns * sizeof(double))) != hipSuccess) return hip_fail(e, "hipMalloc");

  hipLaunchKernelGGL(fill_errors, dim3(1), dim3(kSetupBlock), 0, st, d_err);
  hipLaunchKernelGGL(cell_geometry, dim3(grid_of(Ks)), dim3(kSetupBlock), 0, st, K, num_vertices, d_cx, d_cy,
                     d_etov, d->geo, d->J, d->face_J, d_fJc, d_key, d_val, d_twin, d_err);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "cell_geometry");

  // stable radix sort of the half-edge keys over the bits vertex ids can use
  int vbits = 1;
  while (vbits < 31 && (1u << vbits) < unsigned(num_vertices)) ++vbits;
  size_t sort_bytes = 0, scan_bytes = 0;
  if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, d_key, d_skey, d_val, d_sval, n, 0, 32 + vbits,
                                              st)) != hipSuccess)
    return hip_fail(e, "SortPairs");
  if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, d_flag, d_pos, n, st)) != hipSuccess)
    return hip_fail(e, "ExclusiveSum");
  unsigned char* d_temp;
  if ((e = tmp.alloc(&d_temp, std::max(sort_bytes, scan_bytes))) != hipSuccess) return hip_fail(e, "hipMalloc");
  if ((e = hipcub::DeviceRadixSort::SortPairs(d_temp, sort_bytes, d_key, d_skey, d_val, d_sval, n, 0, 32 + vbits,
                                              st)) != hipSuccess)
    return hip_fail(e, "SortPairs");
  hipLaunchKernelGGL(pair_twins, dim3(grid_of(ns)), dim3(kSetupBlock), 0, st, n, d_skey, d_sval, d_twin, d_err);
  hipLaunchKernelGGL(boundary_flags, dim3(grid_of(ns)), dim3(kSetupBlock), 0, st, n, d_twin, d_flag);
  if ((e = hipcub::DeviceScan::ExclusiveSum(d_temp, scan_bytes, d_flag, d_pos, n, st)) != hipSuccess)
    return hip_fail(e, "ExclusiveSum");

  int errs[kErrCount], tail[2];
  if ((e = hipMemcpyAsync(errs, d_err, sizeof(errs), hipMemcpyDeviceToHost, st)) != hipSuccess)
    return hip_fail(e, "hipMemcpy");
  if ((e = hipMemcpyAsync(&tail[0], d_pos + (n - 1), sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess)
    return hip_fail(e, "hipMemcpy");
  if ((e = hipMemcpyAsync(&tail[1], d_flag + (n - 1), sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess)
    return hip_fail(e, "hipMemcpy");
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "set-up kernels");
  // report in the host path's order: vertex ids, edges, cells
  if (errs[kErrVertex] != INT_MAX)
    return fail(DGC_ERR_INVALID, "dgc_disc2d_create_from_cells: vertex id out of range at cell " +
                                     std::to_string(errs[kErrVertex] / 3));
  if (errs[kErrEdge] != INT_MAX)
    return fail(DGC_ERR_INVALID, "dgc_disc2d_create_from_cells: degenerate edge at cell " +
                                     std::to_string(errs[kErrEdge] / 3));
  if (errs[kErrManifold] != INT_MAX)
    return fail(DGC_ERR_INVALID, "dgc_disc2d_create_from_cells: non-manifold edge");
  if (errs[kErrCell] != INT_MAX)
    return fail(DGC_ERR_INVALID, "dgc_disc2d_create_from_cells: cell " + std::to_string(errs[kErrCell]) +
                                     " is clockwise or degenerate (J <= 0)");
  const long long nbn = (long long)(tail[0] + tail[1]) * Nfp;
  if (nbn >= (1LL << 29)) return fail(DGC_ERR_INVALID, "dgc_disc2d_create_from_cells: too many boundary nodes");
  d->num_boundary_nodes = int(nbn);
  const size_t nb = size_t(nbn > 0 ? nbn : 1);
  if ((e = hipMalloc(&d->bx, nb * sizeof(double))) != hipSuccess) return hip_fail(e, "hipMalloc");
  if ((e = hipMalloc(&d->by, nb * sizeof(double))) != hipSuccess) return hip_fail(e, "hipMalloc");
  hipLaunchKernelGGL(pack_faces, dim3(grid_of(ns)), dim3(kSetupBlock), 0, st, K, Nfp, d_cx, d_cy, d_twin, d_pos,
                     d_frs, d->nbr, d->bx, d->by);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "pack_faces");

  // host copies kept for set-up work that needs them (penalty factors, blocks)
  d->h_J.resize(Ks);
  d->h_face_J.resize(ns);
  d->h_nbr.resize(ns);
  if ((e = hipMemcpyAsync(d->h_J.data(), d->J, Ks * sizeof(double), hipMemcpyDeviceToHost, st)) != hipSuccess)
    return hip_fail(e, "hipMemcpy");
  if ((e = hipMemcpyAsync(d->h_face_J.data(), d_fJc, ns * sizeof(double), hipMemcpyDeviceToHost, st)) != hipSuccess)
    return hip_fail(e, "hipMemcpy");
  if ((e = hipMemcpyAsync(d->h_nbr.data(), d->nbr, ns * sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess)
    return hip_fail(e, "hipMemcpy");
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "set-up kernels");
  *out = d;
  return DGC_OK;
}

int dgc_disc2d_arrays(dgc_disc_t disc, double* geo, int* nbr, double* J, double* face_J, int* num_boundary_nodes,
                      double* bx, double* by) {
  DGC_REQUIRE(disc != nullptr, "dgc_disc2d_arrays: null discretization");
  hipStream_t st = disc->ctx->stream;
  const size_t K = size_t(disc->K);
  DGC_HIP_TRY(hipSetDevice(disc->ctx->device));
  if (geo) DGC_HIP_TRY(hipMemcpyAsync(geo, disc->geo, size_t(dgc::kGeoComps) * K * sizeof(double),
                                      hipMemcpyDeviceToHost, st));
  if (nbr) DGC_HIP_TRY(hipMemcpyAsync(nbr, disc->nbr, 3 * K * sizeof(int), hipMemcpyDeviceToHost, st));
  if (J) DGC_HIP_TRY(hipMemcpyAsync(J, disc->J, K * sizeof(double), hipMemcpyDeviceToHost, st));
  if (face_J) DGC_HIP_TRY(hipMemcpyAsync(face_J, disc->face_J, 3 * K * sizeof(double), hipMemcpyDeviceToHost, st));
  if (num_boundary_nodes) *num_boundary_nodes = disc->num_boundary_nodes;
  const size_t nb = size_t(disc->num_boundary_nodes);
  if (bx && nb) DGC_HIP_TRY(hipMemcpyAsync(bx, disc->bx, nb * sizeof(double), hipMemcpyDeviceToHost, st));
  if (by && nb) DGC_HIP_TRY(hipMemcpyAsync(by, disc->by, nb * sizeof(double), hipMemcpyDeviceToHost, st));
  DGC_HIP_TRY(hipStreamSynchronize(st));
  return DGC_OK;
}

}  // extern "C"
