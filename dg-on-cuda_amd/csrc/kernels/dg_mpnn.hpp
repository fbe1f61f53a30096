// Learned preconditioner: a 3-layer message-passing network over the cells'
// face adjacency (parameters, layouts, launchers). Network definition:
//   x_c   = [r_c / s_r, zb_c / s_z, g_c]   (Np + Np + 4 features, padded to D0)
//   m_c   = bf16( sum_{n in N(c)} h_n )
//   h^l_c = bf16( relu(A_l h^{l-1}_c + B_l m^{l-1}_c + b_l) ),  l = 1..3, width 64
//   y_c   = W_o h^3_c + b_o
//   z     = zb + s_z * c_t * y,  zb = P_BJ r (block-Jacobi)
// bf16 storage for inputs, weights and activations; fp32 accumulation.
#pragma once
#include <cstdint>

#include "dg_common.hpp"

namespace dgc {

constexpr int kHid = 64;        // hidden width
constexpr int kMpnnCells = 128; // cells per workgroup of the layer kernels (4 waves x 32)
constexpr int kPrecondCells = 64;  // cells per workgroup of the preconditioner output

struct MpnnLayout {
  int Np = 0, D0 = 0;  // D0 = 2 Np + 4 rounded up to a multiple of 8
  int oA[3] = {0, 0, 0}, oB[3] = {0, 0, 0}, ob[3] = {0, 0, 0};
  int oWo = 0, obo = 0, total = 0;
  __host__ __device__ int din(int l) const { return l == 0 ? D0 : kHid; }  // width of h^l
};

inline MpnnLayout mpnn_layout(int Np) {
  MpnnLayout L;
  L.Np = Np;
  L.D0 = ((2 * Np + 4) + 7) / 8 * 8;
  int o = 0;
  for (int l = 0; l < 3; ++l) {
    const int d = L.din(l);
    L.oA[l] = o;
    o += kHid * d;
    L.oB[l] = o;
    o += kHid * d;
    L.ob[l] = o;
    o += kHid;
  }
  L.oWo = o;
  o += Np * kHid;
  L.obo = o;
  o += Np;
  L.total = o;
  return L;
}

// Device scalar slots of the network's input/output scales
enum MpnnScalar { kSr = 0, kSz = 1, kCt = 2, kMpnnScalars = 4 };

struct MpnnNet {
  MpnnLayout L;
  int K = 0;
  const int* nbr = nullptr;          // packed face connectivity [3][K]
  const uint16_t* wbf = nullptr;     // bf16 copy of all parameters
  const float* wf = nullptr;         // fp32 master parameters (biases read from here)
  const uint64_t* wpk[3] = {nullptr, nullptr, nullptr};  // MFMA-packed layer weights
};

// forward (cgscal: optional CG scalars; the kernels return at once when its done flag is set)
hipError_t mpnn_features(int order, const MpnnNet& net, const double* r, const double* zb, const double* geo,
                         const float* gnorm, const double* dscal, uint16_t* X, hipStream_t s,
                         const double* cgscal = nullptr);
hipError_t mpnn_dense_fwd(const MpnnNet& net, int layer, const uint16_t* hin, uint16_t* hout, hipStream_t s,
                          const double* cgscal = nullptr);
// out modes
enum MpnnOutMode { kOutY = 0, kOutPrecond = 1, kOutTrain = 2 };
struct MpnnOutArgs {
  const uint16_t* h3 = nullptr;
  float* y = nullptr;           // kOutY: y [K][Np]
  const double* zb = nullptr;   // kOutPrecond: z = zb + s_z c_t y (node-major); partials per 64-cell tile
  double* z = nullptr;
  const double* r = nullptr;    // optional partial dots r.z and z.q
  const double* q = nullptr;
  double* partial = nullptr;    // [2 * grid]
  const double* dscal = nullptr;
  const float* t = nullptr;     // kOutTrain: targets [K][Np]
  float* dy = nullptr;          // dL/dy [K][Np]
  float* dh3 = nullptr;         // [K][64]
  double* loss_partial = nullptr;  // [grid]
  double inv_n = 0.0;           // 1 / (K Np) for the mean loss
};
hipError_t mpnn_out(int order, const MpnnNet& net, int mode, const MpnnOutArgs& a, hipStream_t s,
                    const double* cgscal = nullptr);
// energy-norm loss: z = zb + s y, w = S z -> dy = 2 s (w - r)/(e.r), dh3, partials (z.w, z.r)
hipError_t mpnn_energy_grad(int order, const MpnnNet& net, const double* z, const double* w, const double* r,
                            const double* mscal, double er, float* dy, float* dh3, double* partial, hipStream_t s);
// backward
hipError_t mpnn_dense_bwd(const MpnnNet& net, int layer, const uint16_t* hl, float* dh, float* u, float* v,
                          hipStream_t s);
hipError_t mpnn_nbr_accum(const MpnnNet& net, int din, const float* u, const float* v, float* dhprev, hipStream_t s);
// partial[b][off + o*DX + i] = sum over the block's cells of dz[c][o] * x[c][i]
// (x = hin or its bf16-rounded neighbour sum); bias partial at boff if >= 0.
hipError_t mpnn_wgrad(const MpnnNet& net, int dz_width, const float* dz, int dx_width, const uint16_t* hin,
                      bool gather, float* partial, int pstride, int off, int boff, hipStream_t s);
hipError_t mpnn_reduce(const float* partial, int nblocks, int pstride, int P, float* grad, hipStream_t s);
hipError_t mpnn_adam(float* w, float* m, float* v, const float* g, int P, float lr, float b1, float b2, float eps,
                     int step, uint16_t* wbf, hipStream_t s);
hipError_t mpnn_pack(const MpnnLayout& L, const uint16_t* wbf, uint64_t* const* wpk, hipStream_t s);
int mpnn_grid(int K);

// flexible CG / training data (dg_fcg.hip)
hipError_t fcg_scalars(const double* partial, int nparts, int which, double* scal, double* mscal, double n,
                       double* history, hipStream_t s);
hipError_t grf_eval(const double* x, const double* y, const double* amp, int M, size_t n, double* e, hipStream_t s);
hipError_t mpnn_target(int K, int Np, const double* e, const double* zb, const double* mscal, float* t,
                       hipStream_t s);
// out = a x + b y over n doubles
hipError_t axpby_device(double a, const double* x, double b, const double* y, double* out, size_t n, hipStream_t s);
// out[0] = a . b over n doubles (fixed-order two-level reduction; partial holds >= 256 doubles)
hipError_t dot_device(const double* a, const double* b, size_t n, double* partial, double* out, hipStream_t s);

}  // namespace dgc
