// Flexible-CG scalars and the training-data kernels of the learned
// preconditioner: Gaussian-random-field errors (Dirichlet sine modes with a
// k^-2 amplitude spectrum) and the network's regression targets.
#include <algorithm>

#include "dg_ipdg.hpp"
#include "dg_mpnn.hpp"

namespace dgc {

namespace {

// which 0 / 3: partials (z.z, r.r) after a residual update -> r.r (history,
//   iteration count, done flag) and the network's input scales
//   (3 = initial residual: no iteration increment, history[0]);
// which 1: partials (r.z, z.q) after the network -> rz, beta = -(z.q)/(p.q);
// which 2: partials (r.z, -) at start -> rz, beta = 0.
__global__ void fcg_scalars_kernel(const double* __restrict__ partial, int nparts, int which, double* scal,
                                   double* mscal, double n, double* history) {
  __shared__ double sh[2][1024];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    a += partial[2 * i];
    b += partial[2 * i + 1];
  }
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = b;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (which == 3) {
    const double rr = sh[1][0];
    scal[kRr] = rr;
    if (history) history[0] = rr;
    scal[kDone] = rr <= scal[kTol2] ? 1.0 : 0.0;
    mscal[kSr] = sqrt(rr / n);
    mscal[kSz] = sqrt(sh[0][0] / n);
    return;
  }
  if (scal[kDone] != 0.0) return;
  if (which == 0) {
    const double rr = sh[1][0];
    const int it = int(scal[kIter]);
    scal[kRr] = rr;
    if (history) history[it + 1] = rr;
    scal[kIter] = double(it + 1);
    if (rr <= scal[kTol2] || it + 1 >= int(scal[kMaxit]) || !(rr == rr)) scal[kDone] = 1.0;
    mscal[kSr] = sqrt(rr / n);
    mscal[kSz] = sqrt(sh[0][0] / n);
  } else if (which == 1) {
    scal[kRz] = sh[0][0];
    scal[kBeta] = -sh[1][0] / scal[kPq];
  } else {
    scal[kRz] = sh[0][0];
    scal[kBeta] = 0.0;
  }
}

// e[j*K + c] = sum_{m,n=1..M} amp[(m-1)*M + n-1] sin(m pi (x+1)/2) sin(n pi (y+1)/2)
__global__ __launch_bounds__(256) void grf_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                  const double* __restrict__ amp, int M, size_t n,
                                                  double* __restrict__ e) {
  __shared__ double a[32 * 32];
  for (int i = threadIdx.x; i < M * M; i += blockDim.x) a[i] = amp[i];
  __syncthreads();
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double tx = 0.5 * M_PI * (x[i] + 1.0), ty = 0.5 * M_PI * (y[i] + 1.0);
  double sy[32];
  const double cy = cos(ty);
  sy[0] = sin(ty);
  if (M > 1) sy[1] = 2.0 * cy * sy[0];
  for (int k = 2; k < M; ++k) sy[k] = 2.0 * cy * sy[k - 1] - sy[k - 2];
  const double cx = cos(tx);
  double sm1 = 0.0, sm = sin(tx), acc = 0.0;
  for (int m = 0; m < M; ++m) {
    double row = 0.0;
    for (int k = 0; k < M; ++k) row += a[m * M + k] * sy[k];
    acc += sm * row;
    const double next = 2.0 * cx * sm - sm1;
    sm1 = sm;
    sm = next;
  }
  e[i] = acc;
}

// t[c*Np + j] = (e - zb)[j*K + c] / (s_z c_t)
__global__ void target_kernel(int K, int Np, const double* __restrict__ e, const double* __restrict__ zb,
                              const double* __restrict__ mscal, float* __restrict__ t) {
  const size_t n = size_t(K) * Np;
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = int(i / K), c = int(i % K);
  const double s = mscal[kSz] * mscal[kCt];
  t[size_t(c) * Np + j] = float((e[i] - zb[i]) / s);
}

__global__ __launch_bounds__(256) void dot_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                  size_t n, double* __restrict__ partial) {
  __shared__ double sh[256];
  double s = 0.0;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(256) * gridDim.x) s += a[i] * b[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void sum256_kernel(const double* __restrict__ partial, double* out) {
  __shared__ double sh[256];
  sh[threadIdx.x] = partial[threadIdx.x];
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

__global__ void axpby_dev_kernel(double a, const double* __restrict__ x, double b, const double* __restrict__ y,
                                 double* __restrict__ out, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a * x[i] + b * y[i];
}

}  // namespace

hipError_t axpby_device(double a, const double* x, double b, const double* y, double* out, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(axpby_dev_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, a, x, b, y, out, n);
  return hipGetLastError();
}

hipError_t dot_device(const double* a, const double* b, size_t n, double* partial, double* out, hipStream_t s) {
  hipLaunchKernelGGL(dot_kernel, dim3(256), dim3(256), 0, s, a, b, n, partial);
  hipLaunchKernelGGL(sum256_kernel, dim3(1), dim3(256), 0, s, partial, out);
  return hipGetLastError();
}

hipError_t fcg_scalars(const double* partial, int nparts, int which, double* scal, double* mscal, double n,
                       double* history, hipStream_t s) {
  hipLaunchKernelGGL(fcg_scalars_kernel, dim3(1), dim3(1024), 0, s, partial, nparts, which, scal, mscal, n, history);
  return hipGetLastError();
}

hipError_t grf_eval(const double* x, const double* y, const double* amp, int M, size_t n, double* e, hipStream_t s) {
  if (M < 1 || M > 32) return hipErrorInvalidValue;
  hipLaunchKernelGGL(grf_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, x, y, amp, M, n, e);
  return hipGetLastError();
}

hipError_t mpnn_target(int K, int Np, const double* e, const double* zb, const double* mscal, float* t,
                       hipStream_t s) {
  const size_t n = size_t(K) * Np;
  hipLaunchKernelGGL(target_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, K, Np, e, zb, mscal, t);
  return hipGetLastError();
}

}  // namespace dgc
