// IPDG (BR1-type interior penalty) Laplacian and the CG building blocks.
//
// Replaces src/laplace_2d.h:130-251 (the two-pass apply_homogeneous) and its
// device counterpart examples/poisson_2d/d_poisson_2d.cuh:55-172 launched
// through src/kernel_adapter.cuh:77-82. Boundary data follow the CPU path
// (examples/poisson_2d/poisson_2d.h:73-79: Dirichlet everywhere, u+ = -u-,
// dq.n = 0 at the wall), not the GPU example's affine variant.
//
//   pass 1: q = grad u - LIFT(n [u] / 2),            [u] = u- - u+ (2u- at the wall)
//   pass 2: Lap u = div q - LIFT((n.[q] + 2 tau [u]) / 2), n.[q] = 0 at the wall
//
// Same tiling as the RK kernels: 64 cells per workgroup (lane = cell), node
// rows in LDS, operator rows wave-uniform (scalar loads). Templated on the
// value type so the same code provides the FP64 operator and the FP32 inner
// operator of the mixed-precision solver. Dot products are reduced per
// workgroup into a partial-sum array (fixed order, no atomics) and finished
// by a one-block kernel, so every solve is bit-reproducible.
#include <algorithm>

#include "dg_ipdg.hpp"

namespace dgc {

// Sum over the W waves of a workgroup: wave reduction, then the wave totals
// in wave order; valid in thread 0.
template <int W = kWaves, typename T>
__device__ __forceinline__ T block_sum(T v, T* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[wave] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < W; ++w) s += red[w];
  }
  return s;
}

// The operator tables are invariant while a kernel runs; read through the
// constant address space their wave-uniform rows become scalar loads even
// after the kernel's own global stores (which otherwise force vector loads,
// holding every coefficient in vector registers).
template <typename T>
__device__ __forceinline__ const T __attribute__((address_space(4)))* table(const T* p) {
  return reinterpret_cast<const T __attribute__((address_space(4)))*>(reinterpret_cast<uintptr_t>(p));
}

// Outward normal of face e scaled by face_J / J, from the metric terms of a
// straight-sided cell: -(sx, sy) on face 0, (rx + sx, ry + sy) on face 1,
// -(rx, ry) on face 2. Replaces reading nx, ny and face_J / J per face.
template <typename T>
__device__ __forceinline__ void face_vector(int e, T rx, T ry, T sx, T sy, T& vx, T& vy) {
  const T ax = e == 0 ? sx : (e == 1 ? rx + sx : rx);
  const T ay = e == 0 ? sy : (e == 1 ? ry + sy : ry);
  vx = e == 1 ? ax : -ax;
  vy = e == 1 ? ay : -ay;
}

// W waves serve a 64-cell tile (lane = cell): 4 (256 threads) or 8 (512
// threads, for small meshes: twice the waves in flight, half the rows each).
template <int N, typename T, int W>
__global__ __launch_bounds__(64 * W) void ipdg_pass1(const IpdgParams<T> p) {
  using Tr = Tri<N>;
  constexpr int Np = Tr::Np, Nfp = Tr::Nfp, NF = 3 * Nfp;
  constexpr int RPW = (Np + W - 1) / W;
  constexpr int Npad = ipdg_rows(Np);
  constexpr int FPW = (NF + W - 1) / W;  // face slots per wave (m = wave + W i)
  __shared__ T us[Np][kTile];
  __shared__ T lx[NF][kTile];
  __shared__ T ly[NF][kTile];
  const int K = p.K;
  const int lane = threadIdx.x & (kTile - 1);
  const int wave = uniform(threadIdx.x >> 6);
  const int c0 = blockIdx.x * kTile;
  const int c = c0 + lane;
  const bool valid = c < K;
  const bool fused = p.rec_out != nullptr;
  if (fused) {
    if (p.scal[kDone] != T(0)) {  // finished: carry the record over to the other slot
      if (blockIdx.x == 0 && threadIdx.x < kCgScalars) p.rec_out[threadIdx.x] = p.scal[threadIdx.x];
      return;
    }
  } else if (p.scal && p.scal[kDone] != T(0)) {
    return;
  }
  // With dirz set the operand is the new search direction dirz + beta u.
  auto raw_z = [&](size_t idx) { return p.dirz ? p.dirz[idx] : T(0); };

  // ---- every independent read first: node rows, connectivity, face and
  // metric terms; then the neighbour traces outside the tile
  T own[RPW], ownz[RPW];
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    const int j = wave + W * i;
    const bool ok = valid && j < Np;
    own[i] = ok ? p.u[(size_t)j * K + c] : T(0);
    ownz[i] = ok ? raw_z((size_t)j * K + c) : T(0);
  }
  int code[FPW];
#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wave + W * i;
    code[i] = (valid && m < NF) ? p.nbr[(size_t)(m / Nfp) * K + c] : kBoundaryCode;
  }
  T rx = 0, ry = 0, sx = 0, sy = 0;
  if (valid) {
    rx = p.geo[(size_t)kRx * K + c];
    ry = p.geo[(size_t)kRy * K + c];
    sx = p.geo[(size_t)kSx * K + c];
    sy = p.geo[(size_t)kSy * K + c];
  }
  T far[FPW], farz[FPW];
#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wave + W * i;
    const int d = m % Nfp;
    const int nf = code[i] & 3, idx = code[i] >> 2;
    far[i] = farz[i] = 0;
    if (nf != kBoundaryCode && (unsigned)(idx - c0) >= (unsigned)kTile) {
      const size_t g = (size_t)Tr::fnode(nf, Nfp - 1 - d) * K + idx;
      far[i] = p.u[g];
      farz[i] = raw_z(g);
    }
  }
  // ---- beta: from the record, or (fused) from the update's partials, with
  // the stopping test and the next record
  T beta = 0;
  if (fused) {
    double sums[2];
    sum_partials<2>(p.upart, p.nparts, sums);
    const CgAdvance adv = cg_advance(p.scal, sums[0], sums[1], p.scal[kPhase] != T(0));
    if (blockIdx.x == 0 && threadIdx.x == 0) cg_store(p.scal, p.rec_out, p.history, adv, sums[0], sums[1]);
    if (adv.done) return;
    beta = T(adv.beta);
  } else if (p.dirz) {
    beta = p.scal[kBeta];
  }
  if (p.dirz) {
#pragma unroll
    for (int i = 0; i < RPW; ++i) own[i] = fma(beta, own[i], ownz[i]);
#pragma unroll
    for (int i = 0; i < FPW; ++i) far[i] = fma(beta, far[i], farz[i]);
  }
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    const int j = wave + W * i;
    if (j < Np) {
      us[j][lane] = own[i];
      if (p.dirz && valid) p.dirp[(size_t)j * K + c] = own[i];
    }
  }
  __syncthreads();

#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wave + W * i;
    if (m >= NF) break;
    const int e = m / Nfp, d = m - e * Nfp;
    const T a = us[Tr::fnode(e, d)][lane];
    T du = 0;
    if (valid) {
      const int nf = code[i] & 3, idx = code[i] >> 2;
      if (nf == kBoundaryCode) {
        du = p.neumann ? T(0) : T(2) * a - (p.gbd != nullptr ? T(p.gbd[idx + d]) : T(0));
      } else {
        const int l = idx - c0;
        const T b = ((unsigned)l < (unsigned)kTile) ? us[Tr::fnode(nf, Nfp - 1 - d)][l] : far[i];
        du = a - b;
      }
      p.du[(size_t)m * K + c] = du;
    }
    T vx, vy;  // n face_J / J
    face_vector(e, rx, ry, sx, sy, vx, vy);
    lx[m][lane] = vx * du / T(2);
    ly[m][lane] = vy * du / T(2);
  }
  __syncthreads();

  const int j0 = wave * RPW;
  T dr[RPW], ds[RPW];
#pragma unroll
  for (int i = 0; i < RPW; ++i) dr[i] = ds[i] = 0;
#pragma unroll
  for (int k = 0; k < Np; ++k) {
    const T uk = us[k][lane];
    const auto* a = table(p.DrT) + k * Npad + j0;
    const auto* b = table(p.DsT) + k * Npad + j0;
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      dr[i] = fma(a[i], uk, dr[i]);
      ds[i] = fma(b[i], uk, ds[i]);
    }
  }
  T qx[RPW], qy[RPW];
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    qx[i] = rx * dr[i] + sx * ds[i];
    qy[i] = ry * dr[i] + sy * ds[i];
  }
#pragma unroll
  for (int m = 0; m < NF; ++m) {
    const T fx = lx[m][lane], fy = ly[m][lane];
    const auto* l = table(p.LT) + m * Npad + j0;
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      qx[i] = fma(-l[i], fx, qx[i]);
      qy[i] = fma(-l[i], fy, qy[i]);
    }
  }
  if (!valid) return;
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    const int j = j0 + i;
    if (j >= Np) break;
    p.qx[(size_t)j * K + c] = qx[i];
    p.qy[(size_t)j * K + c] = qy[i];
  }
}

// mode kLaplace: out = Lap u; mode kSpd: out = -J M Lap u (the SPD form),
// optionally with partial[block] = sum over the tile of dotv . out.
template <int N, typename T, int MODE, int W>
__global__ __launch_bounds__(64 * W) void ipdg_pass2(const IpdgParams<T> p) {
  using Tr = Tri<N>;
  constexpr int Np = Tr::Np, Nfp = Tr::Nfp, NF = 3 * Nfp;
  constexpr int RPW = (Np + W - 1) / W;
  constexpr int Npad = ipdg_rows(Np);
  __shared__ T qxs[Np][kTile];
  __shared__ T qys[Np][kTile];
  __shared__ T lf[NF][kTile];
  __shared__ T red[W];
  const int K = p.K;
  const int lane = threadIdx.x & (kTile - 1);
  const int wave = uniform(threadIdx.x >> 6);
  const int c0 = blockIdx.x * kTile;
  const int c = c0 + lane;
  const bool valid = c < K;
  if (p.scal && p.scal[kDone] != T(0)) return;
  constexpr int FPW = (NF + W - 1) / W;  // face slots per wave (m = wave + W i)

  // ---- every independent read first: q rows, connectivity, face terms,
  // penalties, jumps and metric terms; then the neighbours' q outside the tile
  T ownx[RPW], owny[RPW];
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    const int j = wave + W * i;
    const bool ok = valid && j < Np;
    ownx[i] = ok ? p.qx[(size_t)j * K + c] : T(0);
    owny[i] = ok ? p.qy[(size_t)j * K + c] : T(0);
  }
  int code[FPW];
  T ftau[FPW], fdu[FPW];
#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wave + W * i;
    const int e = m / Nfp;
    code[i] = kBoundaryCode;
    ftau[i] = fdu[i] = 0;
    if (valid && m < NF) {
      code[i] = p.nbr[(size_t)e * K + c];
      ftau[i] = p.tau[(size_t)e * K + c];
      fdu[i] = p.du[(size_t)m * K + c];
    }
  }
  T rx = 0, ry = 0, sx = 0, sy = 0;
  if (valid) {
    rx = p.geo[(size_t)kRx * K + c];
    ry = p.geo[(size_t)kRy * K + c];
    sx = p.geo[(size_t)kSx * K + c];
    sy = p.geo[(size_t)kSy * K + c];
  }
  // SPD form: J and the dot operand at this wave's output rows j0 + i
  const T Jc = (MODE == kSpd && valid) ? p.J[c] : T(0);
  T dv[RPW];
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    const int j = wave * RPW + i;
    dv[i] = (MODE == kSpd && p.dotv && valid && j < Np) ? p.dotv[(size_t)j * K + c] : T(0);
  }
  T farx[FPW], fary[FPW];
#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wave + W * i;
    const int d = m % Nfp;
    const int nf = code[i] & 3, idx = code[i] >> 2;
    farx[i] = fary[i] = 0;
    if (nf != kBoundaryCode && (unsigned)(idx - c0) >= (unsigned)kTile) {
      const size_t g = (size_t)Tr::fnode(nf, Nfp - 1 - d) * K + idx;
      farx[i] = p.qx[g];
      fary[i] = p.qy[g];
    }
  }
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    const int j = wave + W * i;
    if (j < Np) {
      qxs[j][lane] = ownx[i];
      qys[j][lane] = owny[i];
    }
  }
  __syncthreads();

#pragma unroll
  for (int i = 0; i < FPW; ++i) {
    const int m = wave + W * i;
    if (m >= NF) break;
    const int e = m / Nfp, d = m - e * Nfp;
    const int own = Tr::fnode(e, d);
    T v = 0;
    if (valid) {
      T vx, vy;  // n face_J / J
      face_vector(e, rx, ry, sx, sy, vx, vy);
      const T fs = sqrt(vx * vx + vy * vy);
      const int nf = code[i] & 3, idx = code[i] >> 2;
      T dqn = 0;
      if (nf != kBoundaryCode) {
        const int nn = Tr::fnode(nf, Nfp - 1 - d);
        const int l = idx - c0;
        T bx, by;
        if ((unsigned)l < (unsigned)kTile) {
          bx = qxs[nn][l];
          by = qys[nn][l];
        } else {
          bx = farx[i];
          by = fary[i];
        }
        dqn = (qxs[own][lane] - bx) * vx + (qys[own][lane] - by) * vy;  // n.[q] face_J / J
      } else if (p.neumann) {
        dqn = T(2) * (qxs[own][lane] * vx + qys[own][lane] * vy) - (p.hbd != nullptr ? T(p.hbd[idx + d]) * fs : T(0));
      }
      v = (dqn + T(2) * (ftau[i] * fdu[i]) * fs) / T(2);
    }
    lf[m][lane] = v;
  }
  __syncthreads();

  const int j0 = wave * RPW;
  // SPD form: the reference mass matrix is folded into the tables
  // (M Dr, M Ds, M LIFT), so -J M Lap u needs no second product
  const T* tDr = MODE == kSpd ? p.MDrT : p.DrT;
  const T* tDs = MODE == kSpd ? p.MDsT : p.DsT;
  const T* tL = MODE == kSpd ? p.MLT : p.LT;
  // div q = Dr (rx qx + ry qy) + Ds (sx qx + sy qy): one accumulator per row
  T lap[RPW];
#pragma unroll
  for (int i = 0; i < RPW; ++i) lap[i] = 0;
#pragma unroll
  for (int k = 0; k < Np; ++k) {
    const T qk = qxs[k][lane], rk = qys[k][lane];
    const T gr = rx * qk + ry * rk, gs = sx * qk + sy * rk;
    const auto* a = table(tDr) + k * Npad + j0;
    const auto* b = table(tDs) + k * Npad + j0;
#pragma unroll
    for (int i = 0; i < RPW; ++i) lap[i] = fma(b[i], gs, fma(a[i], gr, lap[i]));
  }
#pragma unroll
  for (int m = 0; m < NF; ++m) {
    const T f = lf[m][lane];
    const auto* l = table(tL) + m * Npad + j0;
#pragma unroll
    for (int i = 0; i < RPW; ++i) lap[i] = fma(-l[i], f, lap[i]);
  }

  if (MODE == kLaplace) {
    if (valid) {
#pragma unroll
      for (int i = 0; i < RPW; ++i) {
        const int j = j0 + i;
        if (j < Np) p.out[(size_t)j * K + c] = lap[i];
      }
    }
    return;
  }
  // SPD form: out = -J (M Lap u), M already in the tables; with a shift,
  // + shift J M u: the operand's rows replace q_x in LDS (every wave has read
  // q_x above), then the mass rows are applied like the other tables
  if (p.shift != T(0)) {
    const T* sv = p.dirp != nullptr ? p.dirp : p.u;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      const int j = wave + W * i;
      if (j < Np) qxs[j][lane] = valid ? sv[(size_t)j * K + c] : T(0);
    }
    __syncthreads();
    T ms[RPW];
#pragma unroll
    for (int i = 0; i < RPW; ++i) ms[i] = 0;
#pragma unroll
    for (int k = 0; k < Np; ++k) {
      const T v = qxs[k][lane];
      const auto* mm = table(p.MT) + k * Npad + j0;
#pragma unroll
      for (int i = 0; i < RPW; ++i) ms[i] = fma(mm[i], v, ms[i]);
    }
#pragma unroll
    for (int i = 0; i < RPW; ++i) lap[i] = fma(-p.shift, ms[i], lap[i]);
  }
  T dot = 0;
  if (valid) {
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      const int j = j0 + i;
      if (j < Np) {
        const T val = -Jc * lap[i];
        p.out[(size_t)j * K + c] = val;
        dot += dv[i] * val;
      }
    }
  }
  if (p.partial) {
    const T s = block_sum<W>(dot, red);
    if (threadIdx.x == 0) p.partial[blockIdx.x] = s;
  }
}

// out = J M f (the mass-weighted right-hand side)
template <int N, typename T>
__global__ __launch_bounds__(kBlock) void ipdg_mass(const IpdgParams<T> p) {
  using Tr = Tri<N>;
  constexpr int Np = Tr::Np;
  constexpr int RPW = (Np + kWaves - 1) / kWaves;
  constexpr int Npad = ipdg_rows(Np);
  __shared__ T fs[Np][kTile];
  const int K = p.K;
  const int lane = threadIdx.x & (kTile - 1);
  const int wave = uniform(threadIdx.x >> 6);
  const int c = blockIdx.x * kTile + lane;
  const bool valid = c < K;
  for (int j = wave; j < Np; j += kWaves) fs[j][lane] = valid ? p.u[(size_t)j * K + c] : T(0);
  __syncthreads();
  const int j0 = wave * RPW;
  T o[RPW];
#pragma unroll
  for (int i = 0; i < RPW; ++i) o[i] = 0;
#pragma unroll
  for (int k = 0; k < Np; ++k) {
    const T v = fs[k][lane];
    const auto* mm = table(p.MT) + k * Npad + j0;
#pragma unroll
    for (int i = 0; i < RPW; ++i) o[i] = fma(mm[i], v, o[i]);
  }
  if (!valid) return;
  const T Jc = p.J[c];
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    const int j = j0 + i;
    if (j < Np) p.out[(size_t)j * K + c] = Jc * o[i];
  }
}

// CG update with block-Jacobi preconditioning, one tile of cells:
//   x += alpha p;  r -= alpha q;  z = P_c r_c with P_c the cell's symmetric
//   Np x Np inverse block stored packed (entry (j,k), j <= k, of cell c at
//   P[packed_index(Np, j, k) * K + c]; the waves owning rows j and k read the
//   shared entry close together, the second read usually from cache);
//   identity if P is null; partials: r.z (or z.z) and r.r.
template <int N, typename T>
__global__ __launch_bounds__(kBlock) void cg_update(const CgParams<T> p) {
  using Tr = Tri<N>;
  constexpr int Np = Tr::Np;
  constexpr int LPW = (Np + kWaves - 1) / kWaves;  // rows per wave (j = wave + 4 i)
  __shared__ T rs[Np][kTile];
  __shared__ T red[kWaves];
  const int K = p.K;
  const int lane = threadIdx.x & (kTile - 1);
  const int wave = uniform(threadIdx.x >> 6);
  const int c = blockIdx.x * kTile + lane;
  const bool valid = c < K;
  const bool skip = p.scal[kDone] != T(0);  // initialisation: only z and the partials
  if (skip && p.dpart) return;              // fused iteration after the solve finished
  T rv[LPW], xv[LPW], pv[LPW], qv[LPW];
#pragma unroll
  for (int i = 0; i < LPW; ++i) {
    const int j = wave + kWaves * i;
    const bool ok = valid && j < Np;
    const size_t idx = (size_t)j * K + c;
    rv[i] = ok ? p.r[idx] : T(0);
    xv[i] = (ok && !skip) ? p.x[idx] : T(0);
    pv[i] = (ok && !skip) ? p.p[idx] : T(0);
    qv[i] = (ok && !skip) ? p.q[idx] : T(0);
  }
  T alpha;
  if (p.dpart) {
    double pq[1];
    sum_partials<1>(p.dpart, p.nparts, pq);
    alpha = T(double(p.scal[kRz]) / pq[0]);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      p.rec_out[kPq] = T(pq[0]);
      p.rec_out[kAlpha] = alpha;
    }
  } else {
    alpha = p.scal[kAlpha];
  }
  T rr = 0, rz = 0;
#pragma unroll
  for (int i = 0; i < LPW; ++i) {
    const int j = wave + kWaves * i;
    if (j >= Np) break;
    T r = rv[i];
    if (valid && !skip) {
      const size_t idx = (size_t)j * K + c;
      p.x[idx] = xv[i] + alpha * pv[i];
      r = r - alpha * qv[i];
      p.r[idx] = r;
    }
    rs[j][lane] = r;
    rr += r * r;
  }
  __syncthreads();
  for (int j = wave; j < Np; j += kWaves) {
    T z = rs[j][lane];
    if (p.P && valid) {
      z = 0;
#pragma unroll
      for (int k = 0; k < Np; ++k) {
        const int t = k < j ? packed_index(Np, k, j) : packed_index(Np, j, k);
        z = fma(p.P[(size_t)t * K + c], rs[k][lane], z);
      }
    }
    if (valid) p.z[(size_t)j * K + c] = z;
    rz += (p.zz ? z : rs[j][lane]) * z;
  }
  T s1 = block_sum(rz, red);
  __syncthreads();
  T s2 = block_sum(rr, red);
  if (threadIdx.x == 0) {
    p.partial[2 * blockIdx.x] = s1;
    p.partial[2 * blockIdx.x + 1] = s2;
  }
}

// p = z + beta p (beta = 0 on the first iteration)
template <typename T>
__global__ void cg_direction(const T* __restrict__ z, T* __restrict__ pv, const T* __restrict__ scal,
                             size_t n) {
  if (scal[kDone] != T(0)) return;
  const T beta = scal[kBeta];
  const size_t stride = size_t(blockDim.x) * gridDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) pv[i] = z[i] + beta * pv[i];
}

// Finish a reduction of `nparts` partials (the same fixed order as the fused
// iteration's kernels), then update the CG scalars:
//   which == kInit:        rz = r.z, rr = r.r, beta = 0, history[0]; done if rr <= tol2
//   which == kAfterApply:  pq = p.q, alpha = rz / pq
//   which == kAfterUpdate: beta = rz_new / rz, rz = rz_new, rr = r.r,
//                          history[it] = rr, done if rr <= tol2 or it >= maxit
template <typename T>
__global__ __launch_bounds__(kBlock) void cg_scalars(const T* __restrict__ partial, int nparts, int which, T* scal,
                                                     double* history) {
  if (which == kAfterApply) {
    double pq[1];
    sum_partials<1>(partial, nparts, pq);
    if (threadIdx.x != 0 || scal[kDone] != T(0)) return;
    scal[kPq] = T(pq[0]);
    scal[kAlpha] = T(double(scal[kRz]) / pq[0]);
    return;
  }
  double sums[2];
  sum_partials<2>(partial, nparts, sums);
  if (threadIdx.x != 0) return;
  if (which == kAfterUpdate && scal[kDone] != T(0)) return;
  const CgAdvance adv = cg_advance(scal, sums[0], sums[1], which == kInit);
  cg_store(scal, scal, history, adv, sums[0], sums[1]);
}

// ---------------------------------------------------------------------------
namespace {

template <typename T, int N, int W>
void launch_passes(const IpdgParams<T>& p, int mode, int grid, hipStream_t s) {
  if (mode != kPhaseTwo && mode != kSpdPhaseTwo)
    hipLaunchKernelGGL((ipdg_pass1<N, T, W>), dim3(grid), dim3(64 * W), 0, s, p);
  if (mode == kLaplace || mode == kPhaseTwo)
    hipLaunchKernelGGL((ipdg_pass2<N, T, kLaplace, W>), dim3(grid), dim3(64 * W), 0, s, p);
  else if (mode == kSpd || mode == kSpdPhaseTwo)
    hipLaunchKernelGGL((ipdg_pass2<N, T, kSpd, W>), dim3(grid), dim3(64 * W), 0, s, p);
}

template <typename T, int N>
hipError_t launch_apply_n(const IpdgParams<T>& p, int mode, hipStream_t s) {
  const int grid = (p.K + kTile - 1) / kTile;
  if (grid <= kIpdgWideTiles)
    launch_passes<T, N, 8>(p, mode, grid, s);
  else
    launch_passes<T, N, 4>(p, mode, grid, s);
  return hipGetLastError();
}

template <typename T, int N>
hipError_t launch_mass_n(const IpdgParams<T>& p, hipStream_t s) {
  const int grid = (p.K + kTile - 1) / kTile;
  hipLaunchKernelGGL((ipdg_mass<N, T>), dim3(grid), dim3(kBlock), 0, s, p);
  return hipGetLastError();
}

// Inhomogeneous part of the Laplacian (src/laplace_2d.h:253-330,
// apply_inhomogeneous_as_rhs), one thread per cell; cells without a boundary
// face return at once. With g the Dirichlet data (gu, Dirichlet faces) and h
// the Neumann data (hq, Neumann faces) at the boundary nodes of the cell:
//   q     = -(1/J) LIFT (n g face_J / 2)               on the boundary faces
//   inout += Dx qx + Dy qy - (1/J) LIFT ((h + 2 tau g) face_J / 2)
// The reference lifts only the boundary faces in the second product too.
template <int N>
__global__ __launch_bounds__(256) void ipdg_inhomogeneous_kernel(const IpdgParams<double> p, const double* gu,
                                                                  const double* hq, double* inout) {
  using Tr = Tri<N>;
  constexpr int Np = Tr::Np, Nfp = Tr::Nfp, NF = 3 * Nfp, Npad = ipdg_rows(Np);
  const int K = p.K;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= K) return;
  int code[3];
  bool any = false;
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    code[e] = p.nbr[(size_t)e * K + c];
    any |= (code[e] & 3) == kBoundaryCode;
  }
  if (!any) return;
  const double rx = p.geo[(size_t)kRx * K + c], ry = p.geo[(size_t)kRy * K + c];
  const double sx = p.geo[(size_t)kSx * K + c], sy = p.geo[(size_t)kSy * K + c];
  double f[NF];  // lift coefficients of the second product
  double qx[Np], qy[Np];
#pragma unroll
  for (int j = 0; j < Np; ++j) qx[j] = qy[j] = 0.0;
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const bool bnd = (code[e] & 3) == kBoundaryCode;
    const int off = code[e] >> 2;
    double vx, vy;  // n face_J / J
    face_vector(e, rx, ry, sx, sy, vx, vy);
    const double fs = sqrt(vx * vx + vy * vy);
    const double tau = p.tau[(size_t)e * K + c];
#pragma unroll
    for (int d = 0; d < Nfp; ++d) {
      const int m = e * Nfp + d;
      const double g = (bnd && !p.neumann && gu != nullptr) ? gu[off + d] : 0.0;
      const double h = (bnd && p.neumann && hq != nullptr) ? hq[off + d] : 0.0;
      f[m] = bnd ? (h + 2.0 * tau * g) * fs / 2.0 : 0.0;
      const double lx = vx * g / 2.0, ly = vy * g / 2.0;
#pragma unroll
      for (int j = 0; j < Np; ++j) {
        const double l = p.LT[m * Npad + j];
        qx[j] -= l * lx;
        qy[j] -= l * ly;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < Np; ++j) {
    double rqx = 0.0, sqx = 0.0, rqy = 0.0, sqy = 0.0;
#pragma unroll
    for (int k = 0; k < Np; ++k) {
      const double a = p.DrT[k * Npad + j], b = p.DsT[k * Npad + j];
      rqx = fma(a, qx[k], rqx);
      sqx = fma(b, qx[k], sqx);
      rqy = fma(a, qy[k], rqy);
      sqy = fma(b, qy[k], sqy);
    }
    double v = rx * rqx + sx * sqx + ry * rqy + sy * sqy;
#pragma unroll
    for (int m = 0; m < NF; ++m) v -= p.LT[m * Npad + j] * f[m];
    inout[(size_t)j * K + c] += v;
  }
}

template <int N>
hipError_t launch_inhomogeneous_n(const IpdgParams<double>& p, const double* gu, const double* hq, double* inout,
                                  hipStream_t s) {
  hipLaunchKernelGGL((ipdg_inhomogeneous_kernel<N>), dim3((p.K + 255) / 256), dim3(256), 0, s, p, gu, hq, inout);
  return hipGetLastError();
}

template <typename T, int N>
hipError_t launch_update_n(const CgParams<T>& p, hipStream_t s) {
  const int grid = (p.K + kTile - 1) / kTile;
  hipLaunchKernelGGL((cg_update<N, T>), dim3(grid), dim3(kBlock), 0, s, p);
  return hipGetLastError();
}

}  // namespace

#define DGC_ORDER_SWITCH(order, CALL) \
  switch (order) {                    \
    case 1: return CALL(1);           \
    case 2: return CALL(2);           \
    case 3: return CALL(3);           \
    case 4: return CALL(4);           \
    case 5: return CALL(5);           \
    case 6: return CALL(6);           \
    case 7: return CALL(7);           \
    default: return hipErrorInvalidValue; \
  }

template <typename T>
hipError_t ipdg_apply(int order, const IpdgParams<T>& p, int mode, hipStream_t s) {
#define CALL(n) launch_apply_n<T, n>(p, mode, s)
  DGC_ORDER_SWITCH(order, CALL)
#undef CALL
}

template <typename T>
hipError_t ipdg_mass_apply(int order, const IpdgParams<T>& p, hipStream_t s) {
#define CALL(n) launch_mass_n<T, n>(p, s)
  DGC_ORDER_SWITCH(order, CALL)
#undef CALL
}

hipError_t ipdg_inhomogeneous(int order, const IpdgParams<double>& p, const double* gu, const double* hq,
                              double* inout, hipStream_t s) {
#define CALL(n) launch_inhomogeneous_n<n>(p, gu, hq, inout, s)
  DGC_ORDER_SWITCH(order, CALL)
#undef CALL
}

template <typename T>
hipError_t cg_update_launch(int order, const CgParams<T>& p, hipStream_t s) {
#define CALL(n) launch_update_n<T, n>(p, s)
  DGC_ORDER_SWITCH(order, CALL)
#undef CALL
}

template <typename T>
hipError_t cg_direction_launch(const T* z, T* pv, const T* scal, size_t n, hipStream_t s) {
  const unsigned grid = unsigned(std::min<size_t>((n + 255) / 256, 8192));
  hipLaunchKernelGGL((cg_direction<T>), dim3(grid), dim3(256), 0, s, z, pv, scal, n);
  return hipGetLastError();
}

template <typename T>
hipError_t cg_scalars_launch(const T* partial, int nparts, int which, T* scal, double* history,
                             hipStream_t s) {
  hipLaunchKernelGGL((cg_scalars<T>), dim3(1), dim3(kBlock), 0, s, partial, nparts, which, scal, history);
  return hipGetLastError();
}

template hipError_t ipdg_apply<double>(int, const IpdgParams<double>&, int, hipStream_t);
template hipError_t ipdg_apply<float>(int, const IpdgParams<float>&, int, hipStream_t);
template hipError_t ipdg_mass_apply<double>(int, const IpdgParams<double>&, hipStream_t);
template hipError_t ipdg_mass_apply<float>(int, const IpdgParams<float>&, hipStream_t);
template hipError_t cg_update_launch<double>(int, const CgParams<double>&, hipStream_t);
template hipError_t cg_update_launch<float>(int, const CgParams<float>&, hipStream_t);
template hipError_t cg_direction_launch<double>(const double*, double*, const double*, size_t, hipStream_t);
template hipError_t cg_direction_launch<float>(const float*, float*, const float*, size_t, hipStream_t);
template hipError_t cg_scalars_launch<double>(const double*, int, int, double*, double*, hipStream_t);
template hipError_t cg_scalars_launch<float>(const float*, int, int, float*, double*, hipStream_t);

}  // namespace dgc
