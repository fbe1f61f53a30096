// Learned preconditioner kernels (network defined in dg_mpnn.hpp).
//
// Forward layers run on the matrix cores: a wave owns 32 cells and computes
// their 64 hidden units as two 32x32 tiles of v_mfma_f32_32x32x16_bf16 over
// the concatenated [own | neighbour-sum] inputs; the layer's weights are
// pre-packed in the B-operand order of k-steps of 8 and staged in LDS once
// per workgroup.
// Backward and weight-gradient kernels read the same bf16 data and
// accumulate in fp32; weight gradients are reduced per workgroup into a
// partial array and summed in a fixed order (deterministic training).
#include <algorithm>
#include <cmath>

#include "dg_ipdg.hpp"
#include "dg_mfma.hpp"
#include "dg_mpnn.hpp"

namespace dgc {

int mpnn_grid(int K) { return (K + kMpnnCells - 1) / kMpnnCells; }

namespace {

__device__ __forceinline__ bf16x4 as_bf16x4(uint64_t u) {
  bf16x4 r;
  __builtin_memcpy(&r, &u, 8);
  return r;
}

__device__ __forceinline__ void load_bf16x4(const uint16_t* p, float out[4]) {
  uint64_t u = *reinterpret_cast<const uint64_t*>(p);
  for (int v = 0; v < 4; ++v) out[v] = bf2f(uint16_t(u >> (16 * v)));
}

// Parameter tables are invariant while a kernel runs and read at
// wave-uniform addresses: through the constant address space they become
// scalar loads instead of per-lane vector loads held in vector registers.
// bf16 weights are read in pairs (entries 2i, 2i+1; the tables start at even
// offsets of the 256-byte-aligned parameter buffer).
using ConstU32 = const uint32_t __attribute__((address_space(4)));
using ConstF32 = const float __attribute__((address_space(4)));
__device__ __forceinline__ ConstU32* weight_pairs(const uint16_t* w) {
  return reinterpret_cast<ConstU32*>(reinterpret_cast<uintptr_t>(w));
}
__device__ __forceinline__ ConstF32* weight_f32(const float* w) {
  return reinterpret_cast<ConstF32*>(reinterpret_cast<uintptr_t>(w));
}
// Entry i of a bf16 table read as pairs.
__device__ __forceinline__ float weight_at(ConstU32* w, int i) {
  const uint32_t u = w[i >> 1];
  return __uint_as_float((i & 1) ? (u & 0xffff0000u) : (u << 16));
}

template <typename T>
__device__ __forceinline__ T block_sum256(T v, T* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < 4; ++w) s += red[w];
  __syncthreads();
  return s;
}

// ---- features ---------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(256) void features_kernel(MpnnNet net, const double* __restrict__ r,
                                                       const double* __restrict__ zb, const double* __restrict__ geo,
                                                       const float* __restrict__ gnorm, const double* dscal,
                                                       const double* cgscal, uint16_t* __restrict__ X) {
  constexpr int Np = Tri<N>::Np;
  if (cgscal && cgscal[kDone] != 0.0) return;
  const int K = net.K, D0 = net.L.D0;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= K) return;
  const double isr = dscal[kSr] > 0 ? 1.0 / dscal[kSr] : 0.0;
  const double isz = dscal[kSz] > 0 ? 1.0 / dscal[kSz] : 0.0;
  uint16_t* x = X + (size_t)c * D0;
#pragma unroll
  for (int j = 0; j < Np; ++j) {
    x[j] = f2bf(float(r[(size_t)j * K + c] * isr));
    x[Np + j] = f2bf(float(zb[(size_t)j * K + c] * isz));
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) x[2 * Np + k] = f2bf(float(geo[(size_t)k * K + c]) / gnorm[k]);
  for (int k = 2 * Np + 4; k < D0; ++k) x[k] = 0;
}

// ---- dense layer forward (MFMA) ---------------------------------------------
template <int D>
__global__ __launch_bounds__(256) void dense_fwd_kernel(MpnnNet net, int layer, const uint16_t* __restrict__ hin,
                                                        uint16_t* __restrict__ hout, const double* cgscal) {
  // k runs over [own (D) | neighbour sum (D)]. The weights are packed for
  // k-steps of 8 (x8 step q, half h: wl[(2 q + h) 64 + lane]); the products
  // take k-steps of 16 on gfx950's 32x32x16 instruction, whose B operand for
  // lane (j, kq) is the two x8 entries of step 2 s + kq at lanes j and j + 32,
  // and a last k-step of 8 where D is not a multiple of 16 (D = 24).
  constexpr int KS = D / 4;   // x8 k-steps over both halves
  constexpr int S16 = D / 16;  // x16 k-steps per half
  constexpr bool kTail = (D % 16) != 0;
  static_assert(D % 8 == 0, "width must be a multiple of 8");
  __shared__ uint64_t wl[KS * 2 * 64];
  if (cgscal && cgscal[kDone] != 0.0) return;
  const uint64_t* wp = net.wpk[layer];
  for (int i = threadIdx.x; i < KS * 2 * 64; i += 256) wl[i] = wp[i];
  __syncthreads();
  const int K = net.K;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 31, kq = lane >> 5;
  const int cbase = blockIdx.x * kMpnnCells + wave * 32;
  const int c = cbase + row;
  const bool valid = c < K;
  int nb[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int code = valid ? net.nbr[(size_t)e * K + c] : kBoundaryCode;
    nb[e] = (code & 3) != kBoundaryCode ? (code >> 2) : -1;
  }
  // every global read of the wave first (own rows, then the neighbour rows
  // and their bf16-rounded sums), the products after: one round trip.
  // Slot q < 2 S16 holds the 4 values at k = 16 (q / 2) + 8 kq + 4 (q % 2)
  // (the x16 operand is slots 2 s, 2 s + 1); slot 2 S16 the tail's 4 values
  // at k = 16 S16 + 4 kq.
  constexpr int NS = 2 * S16 + (kTail ? 1 : 0);
  uint64_t own[NS], nsum[NS];
  auto koff = [&](int q) { return q < 2 * S16 ? 16 * (q / 2) + 8 * kq + 4 * (q % 2) : 16 * S16 + 4 * kq; };
#pragma unroll
  for (int q = 0; q < NS; ++q)
    own[q] = valid ? *reinterpret_cast<const uint64_t*>(hin + (size_t)c * D + koff(q)) : 0;
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    float sm[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      if (nb[e] >= 0) {
        float t[4];
        load_bf16x4(hin + (size_t)nb[e] * D + koff(q), t);
        for (int v = 0; v < 4; ++v) sm[v] += t[v];
      }
    }
    uint64_t u = 0;
    for (int v = 0; v < 4; ++v) u |= uint64_t(f2bf(sm[v])) << (16 * v);
    nsum[q] = u;
  }
  floatx16 acc0 = {}, acc1 = {};
  const int j = lane & 31;
  // B operand of x16 step s of half `base` (0: own, KS / 2: neighbour sum)
  auto bop = [&](int base, int s16, int h) {
    const int q = (base + 2 * s16 + kq) * 2 + h;
    return as_bf16x8(wl[q * 64 + j], wl[q * 64 + 32 + j]);
  };
#pragma unroll
  for (int s16 = 0; s16 < S16; ++s16) {
    const bf16x8 a = as_bf16x8(own[2 * s16], own[2 * s16 + 1]);
    acc0 = mfma32x32x16bf16(a, bop(0, s16, 0), acc0);
    acc1 = mfma32x32x16bf16(a, bop(0, s16, 1), acc1);
  }
  if constexpr (kTail) {
    const bf16x4 a = as_bf16x4(own[2 * S16]);
    const int q = KS / 2 - 1;  // last x8 step of the own half
    acc0 = mfma32x32x8bf16(a, as_bf16x4(wl[(q * 2 + 0) * 64 + lane]), acc0);
    acc1 = mfma32x32x8bf16(a, as_bf16x4(wl[(q * 2 + 1) * 64 + lane]), acc1);
  }
#pragma unroll
  for (int s16 = 0; s16 < S16; ++s16) {
    const bf16x8 a = as_bf16x8(nsum[2 * s16], nsum[2 * s16 + 1]);
    acc0 = mfma32x32x16bf16(a, bop(KS / 2, s16, 0), acc0);
    acc1 = mfma32x32x16bf16(a, bop(KS / 2, s16, 1), acc1);
  }
  if constexpr (kTail) {
    const bf16x4 a = as_bf16x4(nsum[2 * S16]);
    const int q = KS - 1;  // last x8 step of the neighbour half
    acc0 = mfma32x32x8bf16(a, as_bf16x4(wl[(q * 2 + 0) * 64 + lane]), acc0);
    acc1 = mfma32x32x8bf16(a, as_bf16x4(wl[(q * 2 + 1) * 64 + lane]), acc1);
  }
  ConstF32* bias = weight_f32(net.wf + net.L.ob[layer]);
  const float b0 = bias[j], b1 = bias[32 + j];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int cc = cbase + 8 * (r / 4) + 4 * kq + (r % 4);
    if (cc < K) {
      hout[(size_t)cc * kHid + j] = f2bf(fmaxf(acc0[r] + b0, 0.f));
      hout[(size_t)cc * kHid + 32 + j] = f2bf(fmaxf(acc1[r] + b1, 0.f));
    }
  }
}

// ---- output layer -------------------------------------------------------------
template <int N, int MODE>
__global__ __launch_bounds__(256) void out_kernel(MpnnNet net, MpnnOutArgs a, const double* cgscal) {
  constexpr int Np = Tri<N>::Np;
  __shared__ double red[4];
  if (cgscal && cgscal[kDone] != 0.0) return;
  const int K = net.K;
  const int c = blockIdx.x * 256 + threadIdx.x;
  const bool valid = c < K;
  float h[kHid];
#pragma unroll
  for (int q4 = 0; q4 < kHid / 4; ++q4) {
    float t[4] = {0.f, 0.f, 0.f, 0.f};
    if (valid) load_bf16x4(a.h3 + (size_t)c * kHid + 4 * q4, t);
    for (int v = 0; v < 4; ++v) h[4 * q4 + v] = t[v];
  }
  ConstU32* Wo = weight_pairs(net.wbf + net.L.oWo);
  ConstF32* bo = weight_f32(net.wf + net.L.obo);
  float y[Np];
#pragma unroll
  for (int j = 0; j < Np; ++j) {
    float acc = bo[j];
#pragma unroll
    for (int o = 0; o < kHid; ++o) acc = fmaf(weight_at(Wo, j * kHid + o), h[o], acc);
    y[j] = acc;
  }
  if (MODE == kOutY) {
    if (valid)
      for (int j = 0; j < Np; ++j) a.y[(size_t)c * Np + j] = y[j];
    return;
  }
  // training: mean-square loss, dL/dy and dL/dh3
  double loss = 0.0;
  float dy[Np];
#pragma unroll
  for (int j = 0; j < Np; ++j) {
    const float d = valid ? y[j] - a.t[(size_t)c * Np + j] : 0.f;
    loss += double(d) * double(d);
    dy[j] = float(2.0 * double(d) * a.inv_n);
    if (valid) a.dy[(size_t)c * Np + j] = dy[j];
  }
  if (valid) {
#pragma unroll
    for (int o = 0; o < kHid; ++o) {
      float acc = 0.f;
#pragma unroll
      for (int j = 0; j < Np; ++j) acc = fmaf(weight_at(Wo, j * kHid + o), dy[j], acc);
      a.dh3[(size_t)c * kHid + o] = acc;
    }
  }
  const double s = block_sum256(loss, red);
  if (threadIdx.x == 0) a.loss_partial[blockIdx.x] = s;
}

// Preconditioner output with 64 cells per workgroup (lane = cell) and the Np
// output rows split over the 4 waves: four times the waves of one thread per
// cell. The tile's h3 rows (8 KB) are read once, coalesced, into LDS (row
// stride padded to 17 words against bank conflicts) and shared by the waves.
// z = zb + s_z c_t (W_o h3 + b_o) with the same products as out_kernel;
// partials (r.z, z.q) per 64-cell tile.
template <int N>
__global__ __launch_bounds__(256) void precond_out_kernel(MpnnNet net, MpnnOutArgs a, const double* cgscal) {
  constexpr int Np = Tri<N>::Np;
  constexpr int Q = kHid / 4;  // 64-bit words (4 bf16) per h3 row
  __shared__ double red[4];
  __shared__ uint64_t hs[kPrecondCells][Q + 1];
  if (cgscal && cgscal[kDone] != 0.0) return;
  const int K = net.K;
  const int lane = threadIdx.x & 63;
  const int wave = uniform(threadIdx.x >> 6);
  const int c0 = blockIdx.x * kPrecondCells;
  const int c = c0 + lane;
  const bool valid = c < K;
  for (int i = threadIdx.x; i < kPrecondCells * Q; i += 256) {
    const int cell = i / Q, q = i % Q;
    hs[cell][q] = c0 + cell < K ? reinterpret_cast<const uint64_t*>(a.h3 + (size_t)(c0 + cell) * kHid)[q] : 0;
  }
  __syncthreads();
  float h[kHid];
#pragma unroll
  for (int q4 = 0; q4 < Q; ++q4) {
    const uint64_t u = hs[lane][q4];
#pragma unroll
    for (int v = 0; v < 4; ++v) h[4 * q4 + v] = bf2f(uint16_t(u >> (16 * v)));
  }
  ConstU32* Wo = weight_pairs(net.wbf + net.L.oWo);
  ConstF32* bo = weight_f32(net.wf + net.L.obo);
  const double scale = a.dscal[kSz] * a.dscal[kCt];
  double rz = 0.0, zq = 0.0;
  for (int j = wave; j < Np; j += 4) {
    float acc = bo[j];
#pragma unroll
    for (int o = 0; o < kHid; ++o) acc = fmaf(weight_at(Wo, j * kHid + o), h[o], acc);
    if (valid) {
      const size_t idx = (size_t)j * K + c;
      const double z = a.zb[idx] + scale * double(acc);
      a.z[idx] = z;
      if (a.r) rz += a.r[idx] * z;
      if (a.q) zq += z * a.q[idx];
    }
  }
  if (a.partial) {
    const double s1 = block_sum256(rz, red);
    const double s2 = block_sum256(zq, red);
    if (threadIdx.x == 0) {
      a.partial[2 * blockIdx.x] = s1;
      a.partial[2 * blockIdx.x + 1] = s2;
    }
  }
}

// ---- energy-norm loss gradient ----------------------------------------------
// z = zb + s y, w = S z: dL/dy = s * 2 (w - r) / (e.r); also dL/dh3 and the
// loss terms z.w and z.r per workgroup.
template <int N>
__global__ __launch_bounds__(256) void energy_grad_kernel(MpnnNet net, const double* __restrict__ z,
                                                          const double* __restrict__ w,
                                                          const double* __restrict__ r,
                                                          const double* __restrict__ mscal, double er,
                                                          float* __restrict__ dy, float* __restrict__ dh3,
                                                          double* __restrict__ partial) {
  constexpr int Np = Tri<N>::Np;
  __shared__ double red[4];
  const int K = net.K;
  const int c = blockIdx.x * 256 + threadIdx.x;
  const bool valid = c < K;
  const double g = 2.0 * mscal[kSz] * mscal[kCt] / er;
  ConstU32* Wo = weight_pairs(net.wbf + net.L.oWo);
  double zw = 0.0, zr = 0.0;
  float d[Np];
#pragma unroll
  for (int j = 0; j < Np; ++j) {
    d[j] = 0.f;
    if (valid) {
      const size_t idx = (size_t)j * K + c;
      const double zi = z[idx], wi = w[idx], ri = r[idx];
      zw += zi * wi;
      zr += zi * ri;
      d[j] = float(g * (wi - ri));
      dy[(size_t)c * Np + j] = d[j];
    }
  }
  if (valid) {
#pragma unroll
    for (int o = 0; o < kHid; ++o) {
      float acc = 0.f;
#pragma unroll
      for (int j = 0; j < Np; ++j) acc = fmaf(weight_at(Wo, j * kHid + o), d[j], acc);
      dh3[(size_t)c * kHid + o] = acc;
    }
  }
  const double s1 = block_sum256(zw, red);
  const double s2 = block_sum256(zr, red);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s1;
    partial[2 * blockIdx.x + 1] = s2;
  }
}

// ---- dense layer backward ---------------------------------------------------
// dz = dh * (h > 0) (in place in dh); u = A^T dz, v = B^T dz for the layer below.
// A workgroup owns 64 cells (lane = cell); its four waves split the D outputs
// and read the layer's weights, widened to fp32, from LDS (transposed, so a
// lane's 64 products of one output stream through 16-byte broadcast reads).
// Each u, v entry is the same fp32 multiply-add chain over o = 0..63 as the
// oracle's restatement.
template <int D>
__global__ __launch_bounds__(256) void dense_bwd_kernel(MpnnNet net, int layer, const uint16_t* __restrict__ hl,
                                                        float* __restrict__ dh, float* __restrict__ u,
                                                        float* __restrict__ v) {
  __shared__ float4 at[D][kHid / 4], bt[D][kHid / 4];  // A^T, B^T: [input i][output o]
  const int K = net.K;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + lane;
  const bool valid = c < K;
  if (u) {
    ConstU32* A = weight_pairs(net.wbf + net.L.oA[layer]);
    ConstU32* B = weight_pairs(net.wbf + net.L.oB[layer]);
    float* af = reinterpret_cast<float*>(at);
    float* bfl = reinterpret_cast<float*>(bt);
    for (int idx = threadIdx.x; idx < D * kHid; idx += 256) {
      const int i = idx / kHid, o = idx - i * kHid;
      af[idx] = weight_at(A, o * D + i);
      bfl[idx] = weight_at(B, o * D + i);
    }
  }
  float dz[kHid];
#pragma unroll
  for (int q4 = 0; q4 < kHid / 4; ++q4) {
    float t[4] = {0.f, 0.f, 0.f, 0.f};
    if (valid) load_bf16x4(hl + (size_t)c * kHid + 4 * q4, t);
    for (int w = 0; w < 4; ++w) {
      const int o = 4 * q4 + w;
      const float g = valid ? dh[(size_t)c * kHid + o] : 0.f;
      dz[o] = t[w] > 0.f ? g : 0.f;
    }
  }
  __syncthreads();  // every wave has read dh (and the weights are staged)
  if (valid && wave == 0) {
#pragma unroll
    for (int o = 0; o < kHid; ++o) dh[(size_t)c * kHid + o] = dz[o];
  }
  if (!u) return;  // u is a kernel argument: the whole grid returns here or none of it
  for (int i = wave; i < D; i += 4) {
    float su = 0.f, sv = 0.f;
#pragma unroll
    for (int q = 0; q < kHid / 4; ++q) {
      const float4 a4 = at[i][q], b4 = bt[i][q];
      su = fmaf(a4.x, dz[4 * q + 0], su);
      sv = fmaf(b4.x, dz[4 * q + 0], sv);
      su = fmaf(a4.y, dz[4 * q + 1], su);
      sv = fmaf(b4.y, dz[4 * q + 1], sv);
      su = fmaf(a4.z, dz[4 * q + 2], su);
      sv = fmaf(b4.z, dz[4 * q + 2], sv);
      su = fmaf(a4.w, dz[4 * q + 3], su);
      sv = fmaf(b4.w, dz[4 * q + 3], sv);
    }
    if (valid) {
      u[(size_t)c * D + i] = su;
      v[(size_t)c * D + i] = sv;
    }
  }
}

// dhprev_c = u_c + sum_{n in N(c)} v_n (adjoint of the neighbour sum)
__global__ __launch_bounds__(256) void nbr_accum_kernel(MpnnNet net, int D, const float* __restrict__ u,
                                                        const float* __restrict__ v, float* __restrict__ dhp) {
  const int K = net.K;
  const size_t n = size_t(K) * D;
  const size_t stride = size_t(blockDim.x) * gridDim.x;
  for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n; idx += stride) {
    const int c = int(idx / D), i = int(idx % D);
    float s = u[idx];
    for (int e = 0; e < 3; ++e) {
      const int code = net.nbr[(size_t)e * K + c];
      if ((code & 3) != kBoundaryCode) s += v[(size_t)(code >> 2) * D + i];
    }
    dhp[idx] = s;
  }
}

// ---- weight gradients ---------------------------------------------------------
template <int DZ, int DX>
__global__ __launch_bounds__(256) void wgrad_kernel(MpnnNet net, const float* __restrict__ dz,
                                                    const uint16_t* __restrict__ hin, int gather,
                                                    float* __restrict__ partial, int pstride, int off, int boff) {
  constexpr int CH = 32;
  constexpr int NOUT = DZ * DX;
  constexpr int PER = (NOUT + 255) / 256;
  __shared__ float zs[CH][DZ];
  __shared__ float xs[CH][DX];
  const int K = net.K;
  const int c0 = blockIdx.x * kMpnnCells;
  float acc[PER];
  for (int p = 0; p < PER; ++p) acc[p] = 0.f;
  float bacc = 0.f;
  for (int ch = 0; ch < kMpnnCells; ch += CH) {
    for (int idx = threadIdx.x; idx < CH * DZ; idx += 256) {
      const int cc = idx / DZ, o = idx % DZ, c = c0 + ch + cc;
      zs[cc][o] = c < K ? dz[(size_t)c * DZ + o] : 0.f;
    }
    for (int idx = threadIdx.x; idx < CH * DX; idx += 256) {
      const int cc = idx / DX, i = idx % DX, c = c0 + ch + cc;
      float val = 0.f;
      if (c < K) {
        if (!gather) {
          val = bf2f(hin[(size_t)c * DX + i]);
        } else {
          float s = 0.f;
          for (int e = 0; e < 3; ++e) {
            const int code = net.nbr[(size_t)e * K + c];
            if ((code & 3) != kBoundaryCode) s += bf2f(hin[(size_t)(code >> 2) * DX + i]);
          }
          val = bf2f(f2bf(s));  // the forward pass rounds the neighbour sum to bf16
        }
      }
      xs[cc][i] = val;
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int idx = threadIdx.x + 256 * p;
      if (idx < NOUT) {
        const int o = idx / DX, i = idx % DX;
        float s = acc[p];
        for (int cc = 0; cc < CH; ++cc) s = fmaf(zs[cc][o], xs[cc][i], s);
        acc[p] = s;
      }
    }
    if (boff >= 0 && threadIdx.x < DZ)
      for (int cc = 0; cc < CH; ++cc) bacc += zs[cc][threadIdx.x];
    __syncthreads();
  }
  float* out = partial + (size_t)blockIdx.x * pstride;
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int idx = threadIdx.x + 256 * p;
    if (idx < NOUT) out[off + idx] = acc[p];
  }
  if (boff >= 0 && threadIdx.x < DZ) out[boff + threadIdx.x] = bacc;
}

__global__ void reduce_kernel(const float* __restrict__ partial, int nblocks, int pstride, int P,
                              float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  float s = 0.f;
  for (int b = 0; b < nblocks; ++b) s += partial[(size_t)b * pstride + p];
  grad[p] = s;
}

__global__ void adam_kernel(float* w, float* m, float* v, const float* g, int P, float lr, float b1, float b2,
                            float eps, float c1, float c2, uint16_t* wbf) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const float gi = g[p];
  const float mi = b1 * m[p] + (1.f - b1) * gi;
  const float vi = b2 * v[p] + (1.f - b2) * gi * gi;
  m[p] = mi;
  v[p] = vi;
  const float wi = w[p] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[p] = wi;
  wbf[p] = f2bf(wi);
}

// packed[(ks*2 + tile)*64 + lane] = 4 bf16 of Wcat^T[ks*8 + 4*(lane/32) + v][tile*32 + lane%32]
__global__ void pack_kernel(MpnnLayout L, int layer, const uint16_t* __restrict__ wbf, uint64_t* __restrict__ out) {
  const int D = L.din(layer);
  const int KS = D / 4;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= KS * 2 * 64) return;
  const int lane = idx % 64, tile = (idx / 64) % 2, ks = idx / 128;
  const int o = tile * 32 + lane % 32;
  uint64_t u = 0;
  for (int v = 0; v < 4; ++v) {
    const int k = ks * 8 + 4 * (lane / 32) + v;
    const uint16_t w = k < D ? wbf[L.oA[layer] + o * D + k] : wbf[L.oB[layer] + o * D + (k - D)];
    u |= uint64_t(w) << (16 * v);
  }
  out[idx] = u;
}

}  // namespace

// ---------------------------------------------------------------------------
#define DGC_N_SWITCH(order, CALL) \
  switch (order) {                \
    case 1: CALL(1); break;       \
    case 2: CALL(2); break;       \
    case 3: CALL(3); break;       \
    case 4: CALL(4); break;       \
    case 5: CALL(5); break;       \
    case 6: CALL(6); break;       \
    case 7: CALL(7); break;       \
    default: return hipErrorInvalidValue; \
  }

hipError_t mpnn_features(int order, const MpnnNet& net, const double* r, const double* zb, const double* geo,
                         const float* gnorm, const double* dscal, uint16_t* X, hipStream_t s,
                         const double* cgscal) {
  const int grid = (net.K + 255) / 256;
#define CALL(n) hipLaunchKernelGGL(features_kernel<n>, dim3(grid), dim3(256), 0, s, net, r, zb, geo, gnorm, dscal, cgscal, X)
  DGC_N_SWITCH(order, CALL)
#undef CALL
  return hipGetLastError();
}

hipError_t mpnn_dense_fwd(const MpnnNet& net, int layer, const uint16_t* hin, uint16_t* hout, hipStream_t s,
                          const double* cgscal) {
  const int D = net.L.din(layer);
  const dim3 grid(mpnn_grid(net.K));
  switch (D) {
    case 16: hipLaunchKernelGGL(dense_fwd_kernel<16>, grid, dim3(256), 0, s, net, layer, hin, hout, cgscal); break;
    case 24: hipLaunchKernelGGL(dense_fwd_kernel<24>, grid, dim3(256), 0, s, net, layer, hin, hout, cgscal); break;
    case 40: hipLaunchKernelGGL(dense_fwd_kernel<40>, grid, dim3(256), 0, s, net, layer, hin, hout, cgscal); break;
    case 48: hipLaunchKernelGGL(dense_fwd_kernel<48>, grid, dim3(256), 0, s, net, layer, hin, hout, cgscal); break;
    case 64: hipLaunchKernelGGL(dense_fwd_kernel<64>, grid, dim3(256), 0, s, net, layer, hin, hout, cgscal); break;
    case 80: hipLaunchKernelGGL(dense_fwd_kernel<80>, grid, dim3(256), 0, s, net, layer, hin, hout, cgscal); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t mpnn_out(int order, const MpnnNet& net, int mode, const MpnnOutArgs& a, hipStream_t s,
                    const double* cgscal) {
  const int grid = (net.K + 255) / 256;
  const int pgrid = (net.K + kPrecondCells - 1) / kPrecondCells;
#define CALL(n)                                                                                      \
  if (mode == kOutY)                                                                                 \
    hipLaunchKernelGGL((out_kernel<n, kOutY>), dim3(grid), dim3(256), 0, s, net, a, cgscal);         \
  else if (mode == kOutPrecond)                                                                      \
    hipLaunchKernelGGL((precond_out_kernel<n>), dim3(pgrid), dim3(256), 0, s, net, a, cgscal);       \
  else                                                                                               \
    hipLaunchKernelGGL((out_kernel<n, kOutTrain>), dim3(grid), dim3(256), 0, s, net, a, cgscal)
  DGC_N_SWITCH(order, CALL)
#undef CALL
  return hipGetLastError();
}

hipError_t mpnn_energy_grad(int order, const MpnnNet& net, const double* z, const double* w, const double* r,
                            const double* mscal, double er, float* dy, float* dh3, double* partial, hipStream_t s) {
  const int grid = (net.K + 255) / 256;
#define CALL(n) \
  hipLaunchKernelGGL(energy_grad_kernel<n>, dim3(grid), dim3(256), 0, s, net, z, w, r, mscal, er, dy, dh3, partial)
  DGC_N_SWITCH(order, CALL)
#undef CALL
  return hipGetLastError();
}

hipError_t mpnn_dense_bwd(const MpnnNet& net, int layer, const uint16_t* hl, float* dh, float* u, float* v,
                          hipStream_t s) {
  // layer is the index of h^{layer+1}'s weights; u, v have the width of its input
  const int D = net.L.din(layer);
  const dim3 grid((net.K + 63) / 64);
  switch (D) {
    case 16: hipLaunchKernelGGL(dense_bwd_kernel<16>, grid, dim3(256), 0, s, net, layer, hl, dh, u, v); break;
    case 24: hipLaunchKernelGGL(dense_bwd_kernel<24>, grid, dim3(256), 0, s, net, layer, hl, dh, u, v); break;
    case 40: hipLaunchKernelGGL(dense_bwd_kernel<40>, grid, dim3(256), 0, s, net, layer, hl, dh, u, v); break;
    case 48: hipLaunchKernelGGL(dense_bwd_kernel<48>, grid, dim3(256), 0, s, net, layer, hl, dh, u, v); break;
    case 64: hipLaunchKernelGGL(dense_bwd_kernel<64>, grid, dim3(256), 0, s, net, layer, hl, dh, u, v); break;
    case 80: hipLaunchKernelGGL(dense_bwd_kernel<80>, grid, dim3(256), 0, s, net, layer, hl, dh, u, v); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t mpnn_nbr_accum(const MpnnNet& net, int din, const float* u, const float* v, float* dhprev, hipStream_t s) {
  const size_t n = size_t(net.K) * din;
  const unsigned grid = unsigned(std::min<size_t>((n + 255) / 256, 16384));
  hipLaunchKernelGGL(nbr_accum_kernel, dim3(grid), dim3(256), 0, s, net, din, u, v, dhprev);
  return hipGetLastError();
}

hipError_t mpnn_wgrad(const MpnnNet& net, int dz_width, const float* dz, int dx_width, const uint16_t* hin,
                      bool gather, float* partial, int pstride, int off, int boff, hipStream_t s) {
  const dim3 grid(mpnn_grid(net.K));
  const int g = gather ? 1 : 0;
#define WG(DZ, DX)                                                                                            \
  if (dz_width == DZ && dx_width == DX) {                                                                     \
    hipLaunchKernelGGL((wgrad_kernel<DZ, DX>), grid, dim3(256), 0, s, net, dz, hin, g, partial, pstride, off, \
                       boff);                                                                                 \
    return hipGetLastError();                                                                                 \
  }
  WG(64, 16) WG(64, 24) WG(64, 40) WG(64, 48) WG(64, 64) WG(64, 80)
  WG(3, 64) WG(6, 64) WG(10, 64) WG(15, 64) WG(21, 64) WG(28, 64) WG(36, 64)
#undef WG
  return hipErrorInvalidValue;
}

hipError_t mpnn_reduce(const float* partial, int nblocks, int pstride, int P, float* grad, hipStream_t s) {
  hipLaunchKernelGGL(reduce_kernel, dim3((P + 255) / 256), dim3(256), 0, s, partial, nblocks, pstride, P, grad);
  return hipGetLastError();
}

hipError_t mpnn_adam(float* w, float* m, float* v, const float* g, int P, float lr, float b1, float b2, float eps,
                     int step, uint16_t* wbf, hipStream_t s) {
  const float c1 = 1.f - std::pow(b1, float(step));
  const float c2 = 1.f - std::pow(b2, float(step));
  hipLaunchKernelGGL(adam_kernel, dim3((P + 255) / 256), dim3(256), 0, s, w, m, v, g, P, lr, b1, b2, eps, c1, c2,
                     wbf);
  return hipGetLastError();
}

hipError_t mpnn_pack(const MpnnLayout& L, const uint16_t* wbf, uint64_t* const* wpk, hipStream_t s) {
  for (int l = 0; l < 3; ++l) {
    const int n = L.din(l) / 4 * 2 * 64;
    hipLaunchKernelGGL(pack_kernel, dim3((n + 255) / 256), dim3(256), 0, s, L, l, wbf, wpk[l]);
  }
  return hipGetLastError();
}

}  // namespace dgc
