// The IPDG Laplacian as a sparse matrix (include/dgc_hip.h: dgc_ipdg_assemble):
// the "implicit" laplace_2d the reference's Stokes example asks for
// (examples/stokes_2d/stokes_2d.h:74-76), which it builds column by column with
// one operator application per unit vector (stokes_2d.h:388-422, O(n) applies).
//
// (Lap_h u) on cell c depends on u in the cells within two faces of c (pass 1
// reads the face neighbours, pass 2 the neighbours' q and jumps), so row block
// c of the matrix has the columns of those cells. The cells are coloured so
// that no cell has two cells of one colour within two faces (cells of one
// colour are more than four faces apart); probing with a unit value at node k
// of every cell of one colour then yields, in each cell's rows, the column of
// node k of its unique probed neighbour. Colours x nodes applications in all
// (tens of colours, not the reference's n).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../../include/dgc_hip.h"
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

namespace {

// u = 1 at node k of every cell of colour col, 0 elsewhere (node-major)
__global__ void unit_fill_kernel(const int* color, int col, int k, int K, size_t n, double* u) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const int j = int(i / K), c = int(i - size_t(j) * K);
  u[i] = (j == k && color[c] == col) ? 1.0 : 0.0;
}

// One thread per cell: the response in its rows belongs to node k of the
// unique cell of colour col within two faces (if any), the t-th entry of the
// cell's neighbourhood list.
__global__ void unit_extract_kernel(int K, int Np, int k, int col, const int* color, const int* n2_ptr,
                                    const int* n2_cells, const long long* row_ptr, const double* y,
                                    double* values) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= K) return;
  int t = -1;
  for (int q = n2_ptr[c]; q < n2_ptr[c + 1]; ++q)
    if (color[n2_cells[q]] == col) {
      t = q - n2_ptr[c];
      break;
    }
  if (t < 0) return;
  for (int j = 0; j < Np; ++j) values[row_ptr[size_t(j) * K + c] + size_t(t) * Np + k] = y[size_t(j) * K + c];
}

}  // namespace

extern "C" {

int dgc_ipdg_assemble(dgc_ipdg_t s, long long* nnz, long long* row_ptr, int* col_idx, double* values) {
  DGC_REQUIRE(s != nullptr && nnz != nullptr, "dgc_ipdg_assemble: null argument");
  const dgc_disc_s* d = s->disc;
  const int K = d->K, Np = d->Np;
  const size_t n = size_t(K) * Np;
  // face neighbours and the neighbourhoods within two faces (sorted, with the cell itself)
  std::vector<int> nb(3 * size_t(K), -1);
  for (int c = 0; c < K; ++c)
    for (int e = 0; e < 3; ++e) {
      const int code = d->h_nbr[size_t(e) * K + c];
      if ((code & 3) != dgc::kBoundaryCode) nb[3 * size_t(c) + e] = code >> 2;
    }
  std::vector<int> n2_ptr(K + 1, 0), n2_cells;
  n2_cells.reserve(size_t(K) * 10);
  std::vector<int> tmp;
  for (int c = 0; c < K; ++c) {
    tmp.assign(1, c);
    for (int e = 0; e < 3; ++e) {
      const int a = nb[3 * size_t(c) + e];
      if (a < 0) continue;
      tmp.push_back(a);
      for (int f = 0; f < 3; ++f)
        if (nb[3 * size_t(a) + f] >= 0) tmp.push_back(nb[3 * size_t(a) + f]);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    n2_cells.insert(n2_cells.end(), tmp.begin(), tmp.end());
    n2_ptr[c + 1] = int(n2_cells.size());
  }
  // row (j, c), node-major index j K + c, holds |N2(c)| Np entries ordered by (neighbour, node)
  std::vector<long long> rp(n + 1, 0);
  for (int j = 0; j < Np; ++j)
    for (int c = 0; c < K; ++c) {
      const size_t r = size_t(j) * K + c;
      rp[r + 1] = rp[r] + (long long)(n2_ptr[c + 1] - n2_ptr[c]) * Np;
    }
  *nnz = rp[n];
  if (row_ptr) std::copy(rp.begin(), rp.end(), row_ptr);
  if (col_idx == nullptr && values == nullptr) return DGC_OK;
  DGC_REQUIRE(col_idx && values, "dgc_ipdg_assemble: col_idx and values are both needed");
  for (int j = 0; j < Np; ++j)
    for (int c = 0; c < K; ++c) {
      long long p = rp[size_t(j) * K + c];
      for (int q = n2_ptr[c]; q < n2_ptr[c + 1]; ++q)
        for (int k = 0; k < Np; ++k) col_idx[p++] = k * K + n2_cells[q];
    }
  // colouring: no two cells of one colour in any N2 (greedy, first fit)
  std::vector<int> color(K, -1), mark;
  int ncol = 0;
  for (int c = 0; c < K; ++c) {
    for (int q = n2_ptr[c]; q < n2_ptr[c + 1]; ++q) {
      const int x = n2_cells[q];
      for (int r = n2_ptr[x]; r < n2_ptr[x + 1]; ++r) {
        const int y = n2_cells[r];
        if (color[y] >= 0) {
          if (color[y] >= int(mark.size())) mark.resize(color[y] + 1, -1);
          mark[color[y]] = c;
        }
      }
    }
    int col = 0;
    while (col < int(mark.size()) && mark[col] == c) ++col;
    color[c] = col;
    ncol = std::max(ncol, col + 1);
  }
  // device probing
  hipStream_t st = d->ctx->stream;
  DGC_HIP_TRY(hipSetDevice(d->ctx->device));
  int *dcolor = nullptr, *dptr = nullptr, *dcells = nullptr;
  long long* drp = nullptr;
  double *dvals = nullptr, *u = nullptr;
  auto cleanup = [&]() {
    (void)hipFree(dcolor);
    (void)hipFree(dptr);
    (void)hipFree(dcells);
    (void)hipFree(drp);
    (void)hipFree(dvals);
    (void)hipFree(u);
  };
  hipError_t e = hipMalloc(&dcolor, K * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&dptr, (K + 1) * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&dcells, n2_cells.size() * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&drp, rp.size() * sizeof(long long));
  if (e == hipSuccess) e = hipMalloc(&dvals, size_t(rp[n]) * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&u, n * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(dcolor, color.data(), K * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dptr, n2_ptr.data(), (K + 1) * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dcells, n2_cells.data(), n2_cells.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(drp, rp.data(), rp.size() * sizeof(long long), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemsetAsync(dvals, 0, size_t(rp[n]) * sizeof(double), st);
  for (int col = 0; col < ncol && e == hipSuccess; ++col)
    for (int k = 0; k < Np && e == hipSuccess; ++k) {
      hipLaunchKernelGGL(unit_fill_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, (const int*)dcolor, col,
                         k, K, n, u);
      e = hipGetLastError();
      if (e != hipSuccess) break;
      if (dgc_ipdg_apply(s, u) != DGC_OK) {
        cleanup();
        return DGC_ERR_LAUNCH;
      }
      hipLaunchKernelGGL(unit_extract_kernel, dim3((K + 255) / 256), dim3(256), 0, st, K, Np, k, col,
                         (const int*)dcolor, (const int*)dptr, (const int*)dcells, (const long long*)drp,
                         (const double*)u, dvals);
      e = hipGetLastError();
    }
  if (e == hipSuccess) e = hipMemcpyAsync(values, dvals, size_t(rp[n]) * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  cleanup();
  if (e != hipSuccess) return dgc_set_error(DGC_ERR_HIP, std::string("dgc_ipdg_assemble: ") + hipGetErrorString(e));
  return DGC_OK;
}

}  // extern "C"
