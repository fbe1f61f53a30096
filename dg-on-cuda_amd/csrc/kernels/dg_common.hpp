// Shared device-side definitions for the gfx950 DG kernels.
//
// DOF layout in HBM (node-major, as the reference's GPU path):
//   field f, node j, cell c  ->  u_f[j * K + c]
// so that 64 consecutive cells of one node row form one 512-byte coalesced
// wavefront access. A workgroup owns a tile of 64 consecutive cells
// (lane = cell) and its four waves split the node rows of that tile.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace dgc {

constexpr int kTile = 64;     // cells per workgroup tile (one per lane)
constexpr int kWaves = 4;     // waves per workgroup
constexpr int kBlock = kTile * kWaves;
constexpr int kMaxOrder = 7;

template <int N>
struct Tri {
  static constexpr int Np = (N + 1) * (N + 2) / 2;
  static constexpr int Nfp = N + 1;
  // d-th counter-clockwise node of reference face f (matches the host tables)
  __host__ __device__ static constexpr int fnode(int f, int d) {
    return f == 0 ? d : (f == 1 ? N + d * N - d * (d - 1) / 2 : Np - 1 - d * (d + 3) / 2);
  }
};

// Geometry components, each stored as a [K] array at geo + comp * K.
enum GeoComp {
  kRx = 0, kRy = 1, kSx = 2, kSy = 3,
  kFs0 = 4,   // face_J / J for faces 0..2
  kNx0 = 7,   // outward normal x for faces 0..2
  kNy0 = 10,  // outward normal y for faces 0..2
  kGeoComps = 13
};

// Packed face connectivity (nbr[e * K + c]):
//   interior face: (neighbour cell << 2) | neighbour face
//   boundary face: (offset into boundary-node arrays << 2) | 3
constexpr int kBoundaryCode = 3;

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Bijective workgroup renumbering: consecutive hardware dispatch slots go to
// the 8 XCDs in turn; give XCD x the contiguous range of logical groups
// [x*per + min(x, rem), ...), so neighbouring tiles share that XCD's L2.
__device__ __forceinline__ int xcd_group(int b, int G) {
  constexpr int X = 8;
  const int x = b % X, i = b / X, per = G / X, rem = G % X;
  return x * per + (x < rem ? x : rem) + i;
}

// Accesses of the RK streams that a stage kernel reads or writes once (the
// accumulator, k3, the step's initial state). With NT they carry the
// nontemporal hint: used when the RK working set exceeds the last-level
// cache, so that these streams do not displace the stage input, whose node
// values are read again by the neighbouring cells.
template <bool NT, typename T>
__device__ __forceinline__ T ld_once(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st_once(T* p, T v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

}  // namespace dgc
