// Library-internal state of an IPDG solver handle.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "dg_internal.hpp"
#include "dg_ipdg.hpp"

struct dgc_ipdg_s;

// Device state of one precision level (operators, geometry, CG vectors).
template <typename T>
struct IpdgLevel {
  T* DrT = nullptr;
  T* DsT = nullptr;
  T* LT = nullptr;
  T* MT = nullptr;
  T *MDrT = nullptr, *MDsT = nullptr, *MLT = nullptr;  // M Dr, M Ds, M LIFT
  T* tau_d = nullptr;
  T* geo_d = nullptr;  // FP64 level aliases the discretization's arrays
  T* J_d = nullptr;
  bool own_geo = false;
  T *qx = nullptr, *qy = nullptr, *du = nullptr;
  T *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *x = nullptr;
  T* p2 = nullptr;  // second search-direction buffer (p alternates between p and p2)
  T* partial = nullptr;   // [2 * grid] partial sums
  T* dpartial = nullptr;  // [grid] partials of p.q in the fused iteration
  T* scal = nullptr;
  T* rec = nullptr;  // [2 * kRecStride] the fused iteration's alternating records
  T* P = nullptr;  // block-Jacobi inverses, symmetric, packed [Np(Np+1)/2][K]
  bool use_P = false;
  bool use_coarse = false;  // add the coarse-space correction (FP64 level only)
  hipGraphExec_t graph = nullptr;
  double* graph_hist = nullptr;
  bool graph_P = false;
  bool graph_coarse = false;

  // tables: DrT, DsT, LT, MT, MDrT, MDsT, MLT (transposed and padded)
  int init(const dgc_ipdg_s& s, const std::vector<std::vector<double>>& tables, const std::vector<double>& geo,
           const std::vector<double>& J, const std::vector<double>& tau);
  void release();
  dgc::IpdgParams<T> params(const dgc_ipdg_s& s) const;
};

struct dgc_ipdg_s {
  dgc_disc_s* disc = nullptr;
  int neumann = 0;  // boundary kind of every boundary face: 0 Dirichlet, 1 Neumann
  double shift = 0.0;  // SPD form: -M Lap_h + shift M
  int tau_mode = 0;
  double tau_scale = 1.0;
  int grid = 0;
  std::vector<double> mass;         // reference mass matrix (row-major)
  std::vector<double> diag_blocks;  // S_cc blocks [Np*Np][K] once computed
  int* color = nullptr;
  double* hist = nullptr;  // residual history (squared norms)
  int hist_len = 0;
  double* dot = nullptr;   // [257] partials and result of a device dot product
  IpdgLevel<double> f64;
  IpdgLevel<float> f32;
  // coarse space of the two-level preconditioner (device arrays owned here)
  dgc::CoarseSpace coarse;
  std::vector<void*> coarse_owned;
  std::vector<double> coarse_A;  // Phi^T S Phi (host copy, row-major)
  int coarse_gen = 0;            // bumped whenever arrays that solver graphs capture change
  int op_gen = 0;                // bumped whenever the SPD operator changes (boundary kind, shift)
  // Iterates recorded by the FP64 PCG for the learned preconditioner's training
  // data: with snap set, the first run of a solve copies x after every
  // graph-replayed chunk of iterations into snap[taken * n] while taken < nsnap
  // (a chunk that finished the solve is not kept).
  double* snap = nullptr;
  int nsnap = 0;
  int snap_taken = 0;
  // Solver iterations are replayed from captured graphs (default) or, for
  // kernel tracing, launched one by one (same kernels, same order).
  bool use_graphs = true;
};

// z += Phi A_c^{-1} Phi^T r with partials (r.z, r.r) or (z.z, r.r) into `partial`.
int coarse_correct(dgc_ipdg_s* s, const double* r, double* z, double* partial, int zz, const double* scal);

// ||b - S x||^2 (S x into q, b - S x into r) and v.w, on the device; both wait
// for the value.
int true_residual2(dgc_ipdg_s* s, const double* b, const double* x, double* q, double* r, double* rr);
int dot2(dgc_ipdg_s* s, const double* v, const double* w, double* out);

// FP64 block-Jacobi PCG continuing from the current iterate f64.x.
int cg_init_fp64_restart(dgc_ipdg_s* s, const double* b, double tol2_times_bb, int maxit, double* history);
int cg_run_fp64(dgc_ipdg_s* s, double* history, int maxit, int* iters);
