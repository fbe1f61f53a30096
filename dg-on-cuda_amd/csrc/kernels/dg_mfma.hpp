// MFMA building blocks for gfx950 (wave64) and their operand layouts.
//
// v_mfma_f32_32x32x16_bf16   D[32x32] += A[32x16] * B[16x32]  (gfx950; bf16 in, fp32 acc)
//   lane l holds A[l%32][8*(l/32) + v] and B[8*(l/32) + v][l%32], v = 0..7
//   (8 bf16 in 4 VGPRs); D register r of lane l is D[8*(r/4) + 4*(l/32) + r%4][l%32].
// v_mfma_f32_32x32x8bf16_1k   D[32x32] += A[32x8] * B[8x32]   (bf16 in, fp32 acc)
//   lane l holds A[l%32][4*(l/32) + v] and B[4*(l/32) + v][l%32], v = 0..3
//   (4 bf16 in 2 VGPRs); D layout as above.
// v_mfma_f32_16x16x4f32       D[16x16] += A[16x4] * B[4x16]   (fp32)
//   lane l holds A[l%16][l/16] and B[l/16][l%16]; D register r of lane l is
//   D[4*(l/16) + r][l%16].
// v_mfma_f64_16x16x4f64       the same shapes and A/B layout in fp64, but the
//   result is row-interleaved: D register pair r of lane l is
//   D[4*r + l/16][l%16].
// dgc_selftest_mfma (dg_selftest.hip) checks these layouts on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dgc {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even float -> bf16 bits
__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t b) {
  const uint32_t u = uint32_t(b) << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

__device__ __forceinline__ floatx16 mfma32x32x8bf16(bf16x4 a, bf16x4 b, floatx16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

// 8 bf16 (bit patterns, element v in bits 16 v .. 16 v + 15 of lo, then hi)
// as the operand of the 32x32x16 instruction
__device__ __forceinline__ bf16x8 as_bf16x8(uint64_t lo, uint64_t hi) {
  const uint64_t u[2] = {lo, hi};
  bf16x8 r;
  __builtin_memcpy(&r, u, 16);
  return r;
}

__device__ __forceinline__ floatx16 mfma32x32x16bf16(bf16x8 a, bf16x8 b, floatx16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ floatx4 mfma16x16x4f32(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// v_mfma_f64_16x16x4f64: the fp32 16x16x4 operand layout, rows of the result
// interleaved (register r of lane l holds row 4 r + l / 16)
typedef double doublex4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ doublex4 mfma16x16x4f64(double a, double b, doublex4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

}  // namespace dgc
