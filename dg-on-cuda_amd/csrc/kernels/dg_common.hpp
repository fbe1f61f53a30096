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

}  // namespace dgc
