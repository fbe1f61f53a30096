// Triangle meshes and per-element geometry for the DG hot path.
//
// Topology follows the half-facet structure of
// src/mesh/simple_triangular_mesh_2d.h:151-217 (twin half-edges keyed by the
// sorted vertex pair; a boundary half-edge is its own twin). Geometry follows
// src/mapping_2d.h:101-158 (Jacobian, signed J = area/2, face J = length/2) and
// src/basic_geom_2d.h:110-153 (outward normal). The marshalling into flat
// arrays follows src/simple_discretization_2d.h:84-218.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace dgc_host {

// Deterministic generator for mesh jitter and synthetic data: splitmix64.
struct SplitMix64 {
  uint64_t state;
  explicit SplitMix64(uint64_t seed) : state(seed) {}
  uint64_t next() {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [0, 1) with 53 random bits
  double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Mesh2D {
  int K = 0;                 // cells
  int Nv = 0;                // unique vertices
  std::vector<double> px, py;   // unique vertex coordinates
  std::vector<int> etov;        // [3K] vertex ids, counter-clockwise
  std::vector<double> cx, cy;   // [3K] per-cell vertex coordinates (periodic meshes: unwrapped)
  std::vector<int> nb_cell;     // [3K] twin half-edge cell (boundary: the cell itself)
  std::vector<int> nb_face;     // [3K] twin half-edge local face (boundary: the face itself)
  bool periodic = false;
};

// nx x ny squares over [x0,x1] x [y0,y1], each split along its (i,j)-(i+1,j+1)
// diagonal into cells 2q = (v00, v10, v11) and 2q+1 = (v00, v11, v01), with
// q = i + j*nx. Every vertex draws (dx, dy) = jitter*h*(2U-1) in row-major
// order (j outer, i inner); on a bounded mesh only interior vertices move.
// A periodic mesh identifies x0~x1 and y0~y1 and moves every vertex.
Mesh2D structured_mesh(int nx, int ny, double x0, double x1, double y0, double y1, double jitter,
                       uint64_t seed, bool periodic);

// Mesh from vertex coordinates and connectivity (e.g. a gmsh import).
Mesh2D mesh_from_arrays(int nv, const double* vx, const double* vy, int k, const int* etov);

// Twin half-edge map (fills nb_cell / nb_face).
void build_topology(Mesh2D& m);

// Minimum edge length over all cells.
double min_edge_length(const Mesh2D& m);

// Per-cell geometry in the reference's array conventions.
struct Geometry {
  int K = 0, N = 0, Nfp = 0;
  std::vector<double> inv_jac;    // [4K] (rx, ry, sx, sy) per cell
  std::vector<double> J;          // [K]
  std::vector<double> face_J;     // [3K]
  std::vector<double> nx, ny;     // [3K]
  std::vector<int> iface_cell;    // [3K] neighbour cell (boundary: self)
  std::vector<int> iface_face;    // [3K] neighbour face (boundary: offset into boundary nodes)
  std::vector<double> bnode_x, bnode_y;  // boundary face nodes in iteration order
  int num_boundary_nodes = 0;
};

// Geometry for order N with node positions (r, s) of the reference triangle.
Geometry build_geometry(const Mesh2D& m, int N, const std::vector<double>& r,
                        const std::vector<double>& s);

// Physical coordinates of every DOF, node-major (x[j*K + c]) or cell-major (x[c*Np + j]).
void dof_coordinates(const Mesh2D& m, const std::vector<double>& r, const std::vector<double>& s,
                     bool node_major, std::vector<double>& x, std::vector<double>& y);

// Minimal reader for gmsh MSH 4.1 ASCII files holding a 2D triangle mesh
// (examples/square_domain_*.msh), returning vertices and triangle connectivity.
// Writer of the same format: one node block and one triangle block (tags
// 1..Nv and 1..K, z = 0), full double precision. Periodic meshes are rejected
// (the format has no vertex identification). Returns false with err set on
// failure.
bool write_gmsh41(const Mesh2D& m, const std::string& path, std::string& err);
bool read_gmsh41(const std::string& path, std::vector<double>& vx, std::vector<double>& vy,
                 std::vector<int>& etov, std::string& err);

}  // namespace dgc_host
