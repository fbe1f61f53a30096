// Parameters and launchers of the IPDG operator and CG kernels.
#pragma once
#include "dg_common.hpp"

namespace dgc {

// kLaplace: both passes, out = Lap u; kSpd: both passes, out = -J M Lap u;
// kPhaseOne / kPhaseTwo: one pass only (the reference's two-phase launch).
enum IpdgMode { kLaplace = 0, kSpd = 1, kPhaseOne = 2, kPhaseTwo = 3 };
// CG scalar slots (device array of 8 values of the solver's type)
enum CgScalar {
  kAlpha = 0, kBeta = 1, kRz = 2, kRr = 3, kIter = 4, kDone = 5, kTol2 = 6, kMaxit = 7,
  kPq = 8,  // p.q of the current iteration (flexible CG's beta)
  kCgScalars = 10
};
enum CgStep { kInit = 0, kAfterApply = 1, kAfterUpdate = 2 };

template <typename T>
struct IpdgParams {
  int K;
  const T* __restrict__ DrT;  // strong-form operators, [Np][Npad] transposed
  const T* __restrict__ DsT;
  const T* __restrict__ LT;   // [3Nfp][Npad]
  const T* __restrict__ MT;   // reference mass, [Np][Npad]
  const T* __restrict__ geo;  // [kGeoComps][K]
  const T* __restrict__ J;    // [K]
  const T* __restrict__ tau;  // [3][K] penalty per face
  const int* __restrict__ nbr;
  const T* u;
  T* qx;
  T* qy;
  T* du;    // [3Nfp][K]
  T* out;
  const T* dotv;  // optional: partial dot with out
  T* partial;     // [grid] partial sums (optional)
  const T* scal;  // optional CG scalars: the kernels return at once when scal[kDone] != 0
  // CG search direction formed on the fly (optional): pass 1 applies the
  // operator to u' = dirz + scal[kBeta] u and stores u' in dirp.
  const T* dirz;
  T* dirp;
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
};

// Index of entry (j, k), j <= k, of a symmetric n x n matrix stored as its
// upper triangle row by row.
__host__ __device__ constexpr int packed_index(int n, int j, int k) { return j * n - j * (j - 1) / 2 + (k - j); }

// CG step kAfterUpdate from the summed (r.z, r.r) (thread 0 of the last workgroup).
template <typename T>
__device__ __forceinline__ void cg_after_update(T* scal, double* history, double rz, double rr) {
  scal[kBeta] = T(rz / double(scal[kRz]));
  scal[kRz] = T(rz);
  scal[kRr] = T(rr);
  const int it = int(scal[kIter]);
  if (history) history[it + 1] = rr;  // history[0] holds the initial residual
  scal[kIter] = T(it + 1);
  if (rr <= double(scal[kTol2]) || it + 1 >= int(scal[kMaxit])) scal[kDone] = T(1);
}

template <typename T>
hipError_t ipdg_apply(int order, const IpdgParams<T>& p, int mode, hipStream_t s);
template <typename T>
hipError_t ipdg_mass_apply(int order, const IpdgParams<T>& p, hipStream_t s);
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
  double* rc = nullptr;               // [nc] restricted vector
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
