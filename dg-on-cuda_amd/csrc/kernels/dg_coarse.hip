// Coarse space of the two-level preconditioner for the IPDG operator: three
// linear functions (1, (x - x_a)/h_a, (y - y_a)/h_a) per aggregate of cells,
// the Galerkin matrix A_c = Phi^T S Phi inverted once (Gauss-Jordan, FP64)
// and stored in FP32, and per-iteration restriction / coarse solve /
// prolongation kernels. With block-Jacobi on the cells this is the additive
// two-level method z = P_BJ r + Phi A_c^{-1} Phi^T r. No reference counterpart:
// the reference solves its Poisson system with an assembled sparse QR
// (examples/poisson_2d/main.cpp:132-223).
//
// All sums run in a fixed order (wave shuffles, per-tile partials), so a
// solve is bit-reproducible.
#include <algorithm>

#include "dg_ipdg.hpp"

namespace dgc {

namespace {

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// sum over the 4 waves of a 256-thread block, fixed order; valid in thread 0
__device__ __forceinline__ double tile_sum(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kWaves; ++w) s += red[w];
  return s;
}

__device__ __forceinline__ bool finished(const double* scal) { return scal && scal[kDone] != 0.0; }

// One wave per aggregate: lanes stride over its cells; per cell the node
// moments (sum v, sum rho v, sum sigma v) are mapped to the three basis
// functions by the cell's coefficients.
template <int N>
__global__ __launch_bounds__(256) void restrict_kernel(const CoarseSpace cs, const double* __restrict__ v,
                                                       double* __restrict__ out, int ostride,
                                                       const double* scal) {
  constexpr int Np = Tri<N>::Np;
  if (finished(scal)) return;
  const int lane = threadIdx.x & 63;
  const int a = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (a >= cs.na) return;
  const int K = cs.K;
  double rho[Np], sig[Np];
#pragma unroll
  for (int j = 0; j < Np; ++j) {
    rho[j] = cs.rs[j];
    sig[j] = cs.rs[Np + j];
  }
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  const int end = cs.agg_ptr[a + 1];
  for (int i = cs.agg_ptr[a] + lane; i < end; i += 64) {
    const int c = cs.agg_cells[i];
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
#pragma unroll
    for (int j = 0; j < Np; ++j) {
      const double x = v[(size_t)j * K + c];
      m0 += x;
      m1 += rho[j] * x;
      m2 += sig[j] * x;
    }
    s0 += m0;
    s1 += double(cs.coef[0 * K + c]) * m0 + double(cs.coef[1 * K + c]) * m1 + double(cs.coef[2 * K + c]) * m2;
    s2 += double(cs.coef[3 * K + c]) * m0 + double(cs.coef[4 * K + c]) * m1 + double(cs.coef[5 * K + c]) * m2;
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  if (lane == 0) {
    out[(size_t)(3 * a) * ostride] = s0;
    out[(size_t)(3 * a + 1) * ostride] = s1;
    out[(size_t)(3 * a + 2) * ostride] = s2;
  }
}

// yc = Ainv rc: one wave per row, 16-byte row loads with rc read through the
// cache next to them (rc holds ld entries, the padding zero), four
// independent partial sums per lane, FP64 accumulation in a fixed order.
__global__ __launch_bounds__(256) void solve_kernel(const CoarseSpace cs, const double* scal) {
  if (finished(scal)) return;
  const int nc = cs.nc, ld4 = cs.ld / 4;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= nc) return;
  const float4* a = reinterpret_cast<const float4*>(cs.Ainv + (size_t)row * cs.ld);
  const double2* x = reinterpret_cast<const double2*>(cs.rc);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  int q = lane;
  for (; q + 192 < ld4; q += 256) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = q + 64 * u;
      const float4 w = a[k];
      const double2 x0 = x[2 * k], x1 = x[2 * k + 1];
      acc[u] += double(w.x) * x0.x + double(w.y) * x0.y + double(w.z) * x1.x + double(w.w) * x1.y;
    }
  }
  for (; q < ld4; q += 64) {
    const float4 w = a[q];
    const double2 x0 = x[2 * q], x1 = x[2 * q + 1];
    acc[0] += double(w.x) * x0.x + double(w.y) * x0.y + double(w.z) * x1.x + double(w.w) * x1.y;
  }
  const double sum = wave_sum((acc[0] + acc[1]) + (acc[2] + acc[3]));
  if (lane == 0) cs.yc[row] = sum;
}

// z += Phi yc over a tile of 64 cells (lane = cell, waves split the nodes);
// partials as the CG update kernel writes them.
template <int N>
__global__ __launch_bounds__(256) void prolong_kernel(const CoarseSpace cs, const double* __restrict__ r,
                                                      double* __restrict__ z, double* __restrict__ partial,
                                                      int zz, const double* scal) {
  constexpr int Np = Tri<N>::Np;
  __shared__ double red[kWaves];
  if (finished(scal)) return;
  const int K = cs.K;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int c = blockIdx.x * kTile + lane;
  double rz = 0.0, rr = 0.0;
  if (c < K) {
    const int a = cs.cell_agg[c];
    const double y0 = cs.yc[3 * a], y1 = cs.yc[3 * a + 1], y2 = cs.yc[3 * a + 2];
    float cf[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) cf[k] = cs.coef[k * K + c];
    for (int j = wave; j < Np; j += kWaves) {
      const double rho = cs.rs[j], sig = cs.rs[Np + j];
      const double p1 = double(cf[0]) + double(cf[1]) * rho + double(cf[2]) * sig;
      const double p2 = double(cf[3]) + double(cf[4]) * rho + double(cf[5]) * sig;
      const size_t idx = (size_t)j * K + c;
      const double zn = z[idx] + (y0 + y1 * p1 + y2 * p2);
      z[idx] = zn;
      const double rv = r[idx];
      rr += rv * rv;
      rz += (zz ? zn : rv) * zn;
    }
  }
  const double s1 = tile_sum(rz, red);
  __syncthreads();
  const double s2 = tile_sum(rr, red);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s1;
    partial[2 * blockIdx.x + 1] = s2;
  }
}

template <int N>
__global__ __launch_bounds__(256) void fill_kernel(const CoarseSpace cs, int col, double* __restrict__ u) {
  constexpr int Np = Tri<N>::Np;
  const int a = col / 3, k = col % 3;
  const int i = cs.agg_ptr[a] + blockIdx.x * 256 + threadIdx.x;
  if (i >= cs.agg_ptr[a + 1]) return;
  const int K = cs.K;
  const int c = cs.agg_cells[i];
  for (int j = 0; j < Np; ++j) {
    double v = 1.0;
    if (k > 0) {
      const int o = 3 * (k - 1);
      v = double(cs.coef[o * K + c]) + double(cs.coef[(o + 1) * K + c]) * cs.rs[j] +
          double(cs.coef[(o + 2) * K + c]) * cs.rs[Np + j];
    }
    u[(size_t)j * K + c] = v;
  }
}

// Gauss-Jordan step k: row[j] = A[k][j] / A[k][k] (row[k] = 1 / A[k][k]),
// col[i] = A[i][k] (col[k] = 0).
__global__ void gj_pivot_kernel(const double* __restrict__ A, int n, int k, double* __restrict__ row,
                                double* __restrict__ col) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double piv = A[(size_t)k * n + k];
  row[i] = i == k ? 1.0 / piv : A[(size_t)k * n + i] / piv;
  col[i] = i == k ? 0.0 : A[(size_t)i * n + k];
}

// A[k][:] = row; A[i][j] = (j == k ? 0 : A[i][j]) - col[i] row[j] for i != k
__global__ void gj_update_kernel(double* __restrict__ A, int n, int k, const double* __restrict__ row,
                                 const double* __restrict__ col) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n) return;
  const size_t idx = (size_t)i * n + j;
  if (i == k) {
    A[idx] = row[j];
  } else {
    A[idx] = (j == k ? 0.0 : A[idx]) - col[i] * row[j];
  }
}

__global__ void symmetrize_kernel(const double* __restrict__ A, int n, int ld, float* __restrict__ out) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= ld) return;
  out[(size_t)i * ld + j] = j < n ? float(0.5 * (A[(size_t)i * n + j] + A[(size_t)j * n + i])) : 0.f;
}

}  // namespace

#define DGC_COARSE_SWITCH(order, CALL) \
  switch (order) {                     \
    case 1: CALL(1); break;            \
    case 2: CALL(2); break;            \
    case 3: CALL(3); break;            \
    case 4: CALL(4); break;            \
    case 5: CALL(5); break;            \
    case 6: CALL(6); break;            \
    case 7: CALL(7); break;            \
    default: return hipErrorInvalidValue; \
  }

hipError_t coarse_restrict(int order, const CoarseSpace& cs, const double* v, double* out, int ostride,
                           const double* scal, hipStream_t s) {
  const dim3 grid((cs.na + 3) / 4);
#define CALL(n) hipLaunchKernelGGL(restrict_kernel<n>, grid, dim3(256), 0, s, cs, v, out, ostride, scal)
  DGC_COARSE_SWITCH(order, CALL)
#undef CALL
  return hipGetLastError();
}

hipError_t coarse_solve(const CoarseSpace& cs, const double* scal, hipStream_t s) {
  const int grid = (cs.nc + 3) / 4;
  hipLaunchKernelGGL(solve_kernel, dim3(grid), dim3(256), 0, s, cs, scal);
  return hipGetLastError();
}

hipError_t coarse_prolong(int order, const CoarseSpace& cs, const double* r, double* z, double* partial, int zz,
                          const double* scal, hipStream_t s) {
  const dim3 grid((cs.K + kTile - 1) / kTile);
#define CALL(n) hipLaunchKernelGGL(prolong_kernel<n>, grid, dim3(256), 0, s, cs, r, z, partial, zz, scal)
  DGC_COARSE_SWITCH(order, CALL)
#undef CALL
  return hipGetLastError();
}

hipError_t coarse_fill(int order, const CoarseSpace& cs, int col, double* u, hipStream_t s) {
  // the largest aggregate bounds the grid; blocks past an aggregate's end exit
  const int grid = (cs.K + 255) / 256;
#define CALL(n) hipLaunchKernelGGL(fill_kernel<n>, dim3(grid), dim3(256), 0, s, cs, col, u)
  DGC_COARSE_SWITCH(order, CALL)
#undef CALL
  return hipGetLastError();
}

hipError_t spd_inverse_inplace(double* A, int n, double* work, hipStream_t s) {
  const unsigned gx = unsigned((n + 255) / 256);
  for (int k = 0; k < n; ++k) {
    hipLaunchKernelGGL(gj_pivot_kernel, dim3(gx), dim3(256), 0, s, A, n, k, work, work + n);
    hipLaunchKernelGGL(gj_update_kernel, dim3(gx, unsigned(n)), dim3(256), 0, s, A, n, k, work, work + n);
  }
  return hipGetLastError();
}

hipError_t symmetrize_to_float(const double* A, int n, int ld, float* Ainv, hipStream_t s) {
  hipLaunchKernelGGL(symmetrize_kernel, dim3(unsigned((ld + 255) / 256), unsigned(n)), dim3(256), 0, s, A, n, ld,
                     Ainv);
  return hipGetLastError();
}

}  // namespace dgc
