// Reference-triangle numerics for the nodal DG hot path: orthonormal
// Jacobi/Dubiner bases, Gauss-Lobatto-Legendre points of any order,
// warp-and-blend (alpha-optimised) nodes, and the dense operators the
// kernels consume (Dr, Ds in weak or strong form, LIFT, mass).
//
// Follows the construction of src/reference_triangle.h:45-114,
// src/alpha_optimized_nodes_triangle.h:69-189,
// src/orthonormal_basis_triangle.h:45-68 and src/jacobi_polynomial.h:45-112
// (operator assembly: examples/maxwell_2d/maxwell_2d.h:79-112 for the weak
// form, src/laplace_2d.h:87-112 for the strong form), extended to any order
// (the reference caps N at 6, src/quadrature_rules.h:146).
#pragma once
#include <cstddef>
#include <vector>

namespace dgc_host {

// Row-major dense matrix used only at set-up time.
struct Mat {
  int rows = 0, cols = 0;
  std::vector<double> a;
  Mat() = default;
  Mat(int r, int c, double v = 0.0) : rows(r), cols(c), a(std::size_t(r) * c, v) {}
  double& operator()(int i, int j) { return a[std::size_t(i) * cols + j]; }
  double operator()(int i, int j) const { return a[std::size_t(i) * cols + j]; }
};

Mat matmul(const Mat& x, const Mat& y);
Mat transpose(const Mat& x);
// Gauss-Jordan elimination with partial pivoting; throws on a singular matrix.
Mat inverse(const Mat& x);

// Orthonormal Jacobi polynomial P_n^{(alpha,beta)}(x) on [-1,1] and its derivative.
double jacobi_p(double x, double alpha, double beta, int n);
double grad_jacobi_p(double x, double alpha, double beta, int n);

// npts Gauss-Lobatto-Legendre points on [-1,1] (ascending) and weights.
void lgl_points(int npts, std::vector<double>& x, std::vector<double>& w);

// Warp-and-blend nodes on the reference triangle {r,s >= -1, r+s <= 0}.
// Node n = (level i along s, position j along r), i outer, j inner.
void triangle_nodes(int N, std::vector<double>& r, std::vector<double>& s);

// Index of the d-th node (counter-clockwise) on reference face `face`.
// face 0: s=-1, face 1: r+s=0, face 2: r=-1.
int face_node(int N, int face, int d);

enum OperatorForm { kWeakForm = 0, kStrongForm = 1 };

struct RefTriangle {
  int N = 0, Np = 0, Nfp = 0;
  std::vector<double> r, s;
  std::vector<int> fmask;  // [3][Nfp]
  Mat V, Vr, Vs;           // Vandermonde and its r/s derivatives (Np x Np)
  Mat Mass;                // (V V^T)^{-1}
  Mat MassInv;             // V V^T
  Mat Dr, Ds;              // in the requested form
  Mat LIFT;                // Np x 3Nfp
  RefTriangle(int order, OperatorForm form);
};

}  // namespace dgc_host
