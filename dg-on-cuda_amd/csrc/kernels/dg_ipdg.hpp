// Parameters and launchers of the IPDG operator and CG kernels.
#pragma once
#include "dg_common.hpp"

namespace dgc {

// kLaplace: both passes, out = Lap u; kSpd: both passes, out = -J M Lap u;
// kPhaseOne / kPhaseTwo: one pass only (the reference's two-phase launch);
// kSpdPhaseTwo: the second pass of kSpd alone.
enum IpdgMode { kLaplace = 0, kSpd = 1, kPhaseOne = 2, kPhaseTwo = 3, kSpdPhaseTwo = 4 };
// CG scalar slots (device array of 8 values of the solver's type)
enum CgScalar {
  kAlpha = 0, kBeta = 1, kRz = 2, kRr = 3, kIter = 4, kDone = 5, kTol2 = 6, kMaxit = 7,
  kPq = 8,  // p.q of the current iteration (flexible CG's beta)
  kPhase = 9,  // fused iteration: 1 while the record holds a fresh initialisation
  kCgScalars = 10
};
// The fused PCG iteration keeps two records of CG scalars (stride
// kRecStride), the one an iteration reads and the one it writes, alternating
// with the iteration's parity, so no kernel writes a value that other
// workgroups of the same kernel still read.
constexpr int kRecStride = 16;
// Largest partial-sum count for which the fused iteration's kernels finish
// the reductions themselves (every workgroup sums all partials, from cache).
constexpr int kFusedMaxParts = 4096;
// Row stride of the IPDG operator tables (transposed, zero-padded): a multiple
// of 8, so that 4 or 8 waves per tile can each take a contiguous row block.
__host__ __device__ constexpr int ipdg_rows(int Np) { return 8 * ((Np + 7) / 8); }
// Meshes of at most this many 64-cell tiles run the operator passes with 8
// waves per tile (more waves in flight), larger ones with 4. Measured with
// block-Jacobi PCG: 8 waves 2.4 % faster at 512 tiles (K = 32768, N = 3),
// 4 waves 8-9 % faster at 2048 tiles (K = 131072, N = 2 and 3).
constexpr int kIpdgWideTiles = 1024;
enum CgStep { kInit = 0, kAfterApply = 1, kAfterUpdate = 2 };

template <typename T>
struct IpdgParams {
  int K;
  const T* __restrict__ DrT;  // strong-form operators, [Np][Npad] transposed
  const T* __restrict__ DsT;
  const T* __restrict__ LT;   // [3Nfp][Npad]
  const T* __restrict__ MT;   // reference mass, [Np][Npad]
  const T* __restrict__ MDrT;  // M Dr, M Ds, M LIFT (the SPD form's tables), same layouts
  const T* __restrict__ MDsT;
  const T* __restrict__ MLT;
  const T* __restrict__ geo;  // [kGeoComps][K]
  const T* __restrict__ J;    // [K]
  const T* __restrict__ tau;  // [3][K] penalty per face
  const int* __restrict__ nbr;
  int neumann;  // boundary faces: 0 Dirichlet (u+ = -u-, q+ = q-), 1 Neumann (u+ = u-, q+.n = -q-.n)
  // optional boundary data per boundary node (null: homogeneous): the exterior
  // value at interior 0 on Dirichlet faces (u+ = gbd - u-) and the exterior
  // normal derivative at interior 0 on Neumann faces (q+.n = hbd - q-.n)
  const double* gbd;
  const double* hbd;
  // SPD form only: out = -J M Lap u + shift J M u (0: no shift term). The
  // shifted operand is the vector the operator is applied to: dirp when the
  // search direction is formed on the fly, else u.
  T shift;
  const T* u;
  T* qx;
  T* qy;
  T* du;    // [3Nfp][K]
  T* out;
  const T* dotv;  // optional: partial dot with out
  T* partial;     // [grid] partial sums (optional)
  const T* scal;  // optional CG scalars: the kernels return at once when scal[kDone] != 0
  // CG search direction formed on the fly (optional): pass 1 applies the
  // operator to u' = dirz + beta u and stores u' in dirp, beta = scal[kBeta].
  const T* dirz;
  T* dirp;
  // Fused iteration (optional, pass 1): scal is the record this iteration
  // reads; the kernel sums the update's partials upart[2 i], upart[2 i + 1]
  // (r.z, r.r; i < nparts) itself, takes beta and the stopping test from
  // them, and workgroup 0 writes the next record rec_out and history.
  const T* upart;
  int nparts;
  T* rec_out;
  double* history;
};

template <typename T>
struct CgParams {
  int K;
  T* x;
  T* r;
  const T* p;
  const T* q;
  T* z;
  const T* P;     // block-Jacobi inverses, symmetric, packed: [Np(Np+1)/2][K] or null
  const T* scal;  // CG scalars
  T* partial;     // [2 * grid]: (r.z, r.r), or (z.z, r.r) when zz != 0
  int zz = 0;
  // Fused iteration (optional): the kernel sums pass 2's partials
  // dpart[i] (p.q, i < nparts) itself for alpha = scal[kRz] / p.q, and
  // workgroup 0 stores alpha and p.q in rec_out (the record scal points to).
  const T* dpart = nullptr;
  int nparts = 0;
  T* rec_out = nullptr;
};

// Fixed-order sum of W interleaved series partial[W i + w], i < nparts, in
// double, by the first 256 threads of a workgroup of 256 or more; every
// thread receives the sums. Each of those threads adds a strided subset in
// index order, waves combine by butterfly shuffles (identical in every lane:
// IEEE addition commutes), the four wave totals are added in wave order.
// Every workgroup of any kernel, and the one-workgroup scalar kernel, gets
// bitwise the same sums.
template <int W, typename T>
__device__ __forceinline__ void sum_partials(const T* partial, int nparts, double (&out)[W]) {
  __shared__ double wsum[W][kWaves];
  double a[W];
#pragma unroll
  for (int w = 0; w < W; ++w) a[w] = 0.0;
  if (threadIdx.x < kBlock)
    for (int i = threadIdx.x; i < nparts; i += kBlock)
#pragma unroll
      for (int w = 0; w < W; ++w) a[w] += double(partial[W * i + w]);
#pragma unroll
  for (int w = 0; w < W; ++w)
    for (int off = 32; off > 0; off >>= 1) a[w] += __shfl_xor(a[w], off, 64);
  if ((threadIdx.x & 63) == 0 && threadIdx.x < kBlock)
#pragma unroll
    for (int w = 0; w < W; ++w) wsum[w][threadIdx.x >> 6] = a[w];
  __syncthreads();
#pragma unroll
  for (int w = 0; w < W; ++w) out[w] = ((wsum[w][0] + wsum[w][1]) + wsum[w][2]) + wsum[w][3];
  __syncthreads();
}

// Index of entry (j, k), j <= k, of a symmetric n x n matrix stored as its
// upper triangle row by row.
__host__ __device__ constexpr int packed_index(int n, int j, int k) { return j * n - j * (j - 1) / 2 + (k - j); }

// CG scalars after a residual update from the summed (r.z, r.r); after the
// initialisation (init) beta = 0 and the iteration count stays 0.
struct CgAdvance {
  double beta;
  int it;
  bool done;
};
template <typename T>
__device__ __forceinline__ CgAdvance cg_advance(const T* in, double rz, double rr, bool init) {
  CgAdvance s;
  s.beta = init ? 0.0 : rz / double(in[kRz]);
  s.it = init ? 0 : int(in[kIter]) + 1;
  s.done = rr <= double(in[kTol2]) || (!init && s.it >= int(in[kMaxit]));
  return s;
}
// Store the record after advance s (in and out may be the same record);
// history[it] = r.r (history[0]: the initial residual).
template <typename T>
__device__ __forceinline__ void cg_store(const T* in, T* out, double* history, const CgAdvance& s, double rz,
                                         double rr) {
  const T alpha = in[kAlpha], tol2 = in[kTol2], maxit = in[kMaxit], pq = in[kPq];
  out[kAlpha] = alpha;
  out[kBeta] = T(s.beta);
  out[kRz] = T(rz);
  out[kRr] = T(rr);
  out[kIter] = T(s.it);
  out[kDone] = T(s.done ? 1 : 0);
  out[kTol2] = tol2;
  out[kMaxit] = maxit;
  out[kPq] = pq;
  out[kPhase] = T(0);
  if (history) history[s.it] = rr;
}

template <typename T>
hipError_t ipdg_apply(int order, const IpdgParams<T>& p, int mode, hipStream_t s);
template <typename T>
hipError_t ipdg_mass_apply(int order, const IpdgParams<T>& p, hipStream_t s);
// inout += the inhomogeneous part of the Laplacian for boundary data gu
// (Dirichlet exterior values at interior value 0) or hq (Neumann exterior
// normal derivatives at 0), per boundary node; FP64 only.
hipError_t ipdg_inhomogeneous(int order, const IpdgParams<double>& p, const double* gu, const double* hq,
                              double* inout, hipStream_t s);
template <typename T>
hipError_t cg_update_launch(int order, const CgParams<T>& p, hipStream_t s);
template <typename T>
hipError_t cg_direction_launch(const T* z, T* pv, const T* scal, size_t n, hipStream_t s);
template <typename T>
hipError_t cg_scalars_launch(const T* partial, int nparts, int which, T* scal, double* history,
                             hipStream_t s);

// ---- coarse space (dg_coarse.hip) -------------------------------------------
// Three linear functions per aggregate of cells, phi_0 = 1,
// phi_1 = (x - x_a)/h_a, phi_2 = (y - y_a)/h_a. On cell c they are affine in the
// reference coordinates (rho, sigma) of its nodes:
//   phi_k(node j) = coef[3k-3][c] + coef[3k-2][c] rho_j + coef[3k-1][c] sigma_j  (k = 1, 2).
struct CoarseSpace {
  int K = 0, na = 0, nc = 0, ld = 0;  // cells, aggregates, unknowns (3 na), row stride of Ainv
  const int* agg_ptr = nullptr;       // [na + 1] CSR over the cells of each aggregate
  const int* agg_cells = nullptr;     // [K]
  const int* cell_agg = nullptr;      // [K]
  const float* coef = nullptr;        // [6][K]
  const double* rs = nullptr;         // [2][Np] reference node coordinates (rho, sigma)
  const float* Ainv = nullptr;        // [nc][ld] inverse of Phi^T S Phi (fp32, symmetric)
  double* rc = nullptr;               // [ld] restricted vector (entries nc..ld-1 stay zero)
  double* yc = nullptr;               // [nc] coarse solution
};
// out[(3a + k) * ostride] = sum over the aggregate's nodes of phi_k v (wave per aggregate)
hipError_t coarse_restrict(int order, const CoarseSpace& cs, const double* v, double* out, int ostride,
                           const double* scal, hipStream_t s);
// yc = Ainv rc
hipError_t coarse_solve(const CoarseSpace& cs, const double* scal, hipStream_t s);
// z += Phi yc; partials (r.z, r.r), or (z.z, r.r) when zz != 0, per tile of 64 cells
hipError_t coarse_prolong(int order, const CoarseSpace& cs, const double* r, double* z, double* partial, int zz,
                          const double* scal, hipStream_t s);
// u = column `col` of Phi (u zeroed by the caller)
hipError_t coarse_fill(int order, const CoarseSpace& cs, int col, double* u, hipStream_t s);
// In-place Gauss-Jordan inverse of an SPD n x n matrix (no pivoting); work holds 2n doubles.
hipError_t spd_inverse_inplace(double* A, int n, double* work, hipStream_t s);
// Ainv[i][j] = float((A[i][j] + A[j][i]) / 2), row stride ld, zero padding
hipError_t symmetrize_to_float(const double* A, int n, int ld, float* Ainv, hipStream_t s);

}  // namespace dgc
