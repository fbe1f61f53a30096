// Device self-test of the MFMA operand layouts documented in dg_mfma.hpp.
// One wave multiplies caller-given matrices with the builtin and scatters
// the result with the documented layout; the caller compares with A*B.
#include <string>

#include "../../../include/dgc_hip.h"
#include "dg_internal.hpp"
#include "dg_mfma.hpp"

namespace {

__global__ void mfma_bf16_probe(const float* A, const float* B, float* C) {
  // A: 32x8 row-major, B: 8x32 row-major, C: 32x32 row-major
  const int l = threadIdx.x;
  dgc::bf16x4 a, b;
  for (int v = 0; v < 4; ++v) {
    a[v] = short(dgc::f2bf(A[(l % 32) * 8 + 4 * (l / 32) + v]));
    b[v] = short(dgc::f2bf(B[(4 * (l / 32) + v) * 32 + l % 32]));
  }
  dgc::floatx16 d = {};
  d = dgc::mfma32x32x8bf16(a, b, d);
  for (int r = 0; r < 16; ++r) C[(8 * (r / 4) + 4 * (l / 32) + r % 4) * 32 + l % 32] = d[r];
}

__global__ void mfma_bf16x16_probe(const float* A, const float* B, float* C) {
  // A: 32x16 row-major, B: 16x32 row-major, C: 32x32 row-major
  const int l = threadIdx.x;
  uint64_t a[2] = {0, 0}, b[2] = {0, 0};
  for (int v = 0; v < 8; ++v) {
    a[v / 4] |= uint64_t(dgc::f2bf(A[(l % 32) * 16 + 8 * (l / 32) + v])) << (16 * (v % 4));
    b[v / 4] |= uint64_t(dgc::f2bf(B[(8 * (l / 32) + v) * 32 + l % 32])) << (16 * (v % 4));
  }
  dgc::floatx16 d = {};
  d = dgc::mfma32x32x16bf16(dgc::as_bf16x8(a[0], a[1]), dgc::as_bf16x8(b[0], b[1]), d);
  for (int r = 0; r < 16; ++r) C[(8 * (r / 4) + 4 * (l / 32) + r % 4) * 32 + l % 32] = d[r];
}

__global__ void mfma_f32_probe(const float* A, const float* B, float* C) {
  // A: 16x4, B: 4x16, C: 16x16, row-major
  const int l = threadIdx.x;
  dgc::floatx4 d = {};
  d = dgc::mfma16x16x4f32(A[(l % 16) * 4 + l / 16], B[(l / 16) * 16 + l % 16], d);
  for (int r = 0; r < 4; ++r) C[(4 * (l / 16) + r) * 16 + l % 16] = d[r];
}

__global__ void mfma_f64_probe(const float* A, const float* B, float* C) {
  // A: 16x4, B: 4x16 row-major; products and sums in fp64. C receives the raw
  // result registers, C[4 l + r] = register r of lane l, for the caller to
  // match against the product.
  const int l = threadIdx.x;
  dgc::doublex4 d = {0.0, 0.0, 0.0, 0.0};
  d = dgc::mfma16x16x4f64(double(A[(l % 16) * 4 + l / 16]), double(B[(l / 16) * 16 + l % 16]), d);
  for (int r = 0; r < 4; ++r) C[4 * l + r] = float(d[r]);
}

}  // namespace

extern "C" int dgc_selftest_mfma(dgc_ctx_t ctx, int kind, const float* A, const float* B, float* C) {
  if (!ctx || !A || !B || !C || kind < 0 || kind > 3)
    return dgc_set_error(DGC_ERR_INVALID, "dgc_selftest_mfma: bad argument");
  const int na = kind == 0 ? 32 * 8 : (kind == 3 ? 32 * 16 : 16 * 4);
  const int nb = kind == 0 ? 8 * 32 : (kind == 3 ? 16 * 32 : 4 * 16);
  const int nc = (kind == 0 || kind == 3) ? 32 * 32 : 16 * 16;
  float *dA = nullptr, *dB = nullptr, *dC = nullptr;
  hipError_t e = hipMalloc(&dA, na * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dB, nb * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dC, nc * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(dA, A, na * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dB, B, nb * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(dC, 0, nc * sizeof(float));
  if (e == hipSuccess) {
    if (kind == 0)
      hipLaunchKernelGGL(mfma_bf16_probe, dim3(1), dim3(64), 0, ctx->stream, dA, dB, dC);
    else if (kind == 1)
      hipLaunchKernelGGL(mfma_f32_probe, dim3(1), dim3(64), 0, ctx->stream, dA, dB, dC);
    else if (kind == 3)
      hipLaunchKernelGGL(mfma_bf16x16_probe, dim3(1), dim3(64), 0, ctx->stream, dA, dB, dC);
    else
      hipLaunchKernelGGL(mfma_f64_probe, dim3(1), dim3(64), 0, ctx->stream, dA, dB, dC);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(C, dC, nc * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  if (e != hipSuccess) return dgc_set_error(DGC_ERR_HIP, std::string("dgc_selftest_mfma: ") + hipGetErrorString(e));
  return DGC_OK;
}
