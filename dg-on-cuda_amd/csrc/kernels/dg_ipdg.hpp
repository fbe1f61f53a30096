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
};

template <typename T>
struct CgParams {
  int K;
  T* x;
  T* r;
  const T* p;
  const T* q;
  T* z;
  const T* P;     // block-Jacobi inverse blocks [Np*Np][K] or null
  const T* scal;  // CG scalars
  T* partial;     // [2 * grid]: (r.z, r.r), or (z.z, r.r) when zz != 0
  int zz = 0;
};

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

}  // namespace dgc
