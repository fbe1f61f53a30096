// Reference-triangle numerics (set-up only; nothing here runs per time step).
#include "refelem.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace dgc_host {

Mat matmul(const Mat& x, const Mat& y) {
  if (x.cols != y.rows) throw std::invalid_argument("matmul: shape mismatch");
  Mat z(x.rows, y.cols);
  for (int i = 0; i < x.rows; ++i)
    for (int j = 0; j < y.cols; ++j) {
      double acc = 0.0;
      for (int k = 0; k < x.cols; ++k) acc += x(i, k) * y(k, j);
      z(i, j) = acc;
    }
  return z;
}

Mat transpose(const Mat& x) {
  Mat z(x.cols, x.rows);
  for (int i = 0; i < x.rows; ++i)
    for (int j = 0; j < x.cols; ++j) z(j, i) = x(i, j);
  return z;
}

Mat inverse(const Mat& x) {
  if (x.rows != x.cols || x.rows == 0) throw std::invalid_argument("inverse: not square");
  const int n = x.rows;
  Mat work = x;
  Mat inv(n, n);
  for (int i = 0; i < n; ++i) inv(i, i) = 1.0;
  for (int col = 0; col < n; ++col) {
    int piv = col;
    for (int i = col + 1; i < n; ++i)
      if (std::fabs(work(i, col)) > std::fabs(work(piv, col))) piv = i;
    if (work(piv, col) == 0.0) throw std::runtime_error("inverse: singular matrix");
    if (piv != col)
      for (int j = 0; j < n; ++j) {
        std::swap(work(col, j), work(piv, j));
        std::swap(inv(col, j), inv(piv, j));
      }
    const double d = work(col, col);
    for (int j = 0; j < n; ++j) {
      work(col, j) /= d;
      inv(col, j) /= d;
    }
    for (int i = 0; i < n; ++i) {
      if (i == col) continue;
      const double f = work(i, col);
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) {
        work(i, j) -= f * work(col, j);
        inv(i, j) -= f * inv(col, j);
      }
    }
  }
  return inv;
}

// Three-term recurrence for orthonormal Jacobi polynomials
// (Hesthaven & Warburton, "Nodal DG Methods", App. A, JacobiP).
double jacobi_p(double x, double alpha, double beta, int n) {
  const double ab = alpha + beta;
  const double g0 = std::pow(2.0, ab + 1.0) / (ab + 1.0) * std::tgamma(alpha + 1.0) *
                    std::tgamma(beta + 1.0) / std::tgamma(ab + 1.0);
  double p_prev = 1.0 / std::sqrt(g0);
  if (n == 0) return p_prev;
  const double g1 = (alpha + 1.0) * (beta + 1.0) / (ab + 3.0) * g0;
  double p = ((ab + 2.0) * x / 2.0 + (alpha - beta) / 2.0) / std::sqrt(g1);
  if (n == 1) return p;
  double a_old = 2.0 / (2.0 + ab) * std::sqrt((alpha + 1.0) * (beta + 1.0) / (ab + 3.0));
  for (int i = 1; i < n; ++i) {
    const double h1 = 2.0 * i + ab;
    const double a_new = 2.0 / (h1 + 2.0) *
                         std::sqrt((i + 1.0) * (i + 1.0 + ab) * (i + 1.0 + alpha) *
                                   (i + 1.0 + beta) / (h1 + 1.0) / (h1 + 3.0));
    const double b_new = -(alpha * alpha - beta * beta) / h1 / (h1 + 2.0);
    const double p_next = (-a_old * p_prev + (x - b_new) * p) / a_new;
    a_old = a_new;
    p_prev = p;
    p = p_next;
  }
  return p;
}

double grad_jacobi_p(double x, double alpha, double beta, int n) {
  if (n == 0) return 0.0;
  return std::sqrt(n * (n + alpha + beta + 1.0)) * jacobi_p(x, alpha + 1.0, beta + 1.0, n - 1);
}

// Newton iteration on (1-x^2) P_N'(x) = 0 started from Chebyshev-Gauss-Lobatto
// points, using the Legendre recurrence (von Winckel's lglnodes algorithm).
void lgl_points(int npts, std::vector<double>& x, std::vector<double>& w) {
  if (npts < 2) throw std::invalid_argument("lgl_points: need at least 2 points");
  const int N = npts - 1;
  std::vector<double> xi(npts), xold(npts);
  for (int i = 0; i <= N; ++i) xi[i] = std::cos(M_PI * i / N);
  std::vector<double> PN(npts), PNm1(npts);
  for (int it = 0; it < 100; ++it) {
    double change = 0.0;
    for (int i = 0; i <= N; ++i) {
      double p0 = 1.0, p1 = xi[i];
      for (int k = 2; k <= N; ++k) {
        const double p2 = ((2.0 * k - 1.0) * xi[i] * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      PN[i] = p1;
      PNm1[i] = p0;
      xold[i] = xi[i];
      xi[i] = xold[i] - (xi[i] * PN[i] - PNm1[i]) / (npts * PN[i]);
      change = std::max(change, std::fabs(xi[i] - xold[i]));
    }
    if (change <= 4.0 * std::numeric_limits<double>::epsilon()) break;
  }
  x.assign(npts, 0.0);
  w.assign(npts, 0.0);
  for (int i = 0; i <= N; ++i) {
    double p0 = 1.0, p1 = xi[i];
    for (int k = 2; k <= N; ++k) {
      const double p2 = ((2.0 * k - 1.0) * xi[i] * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    // ascending order: reverse the cos ordering
    x[N - i] = xi[i];
    w[N - i] = 2.0 / (N * npts * p1 * p1);
  }
  x[0] = -1.0;
  x[N] = 1.0;
  // Symmetrise to remove the last-bit asymmetry of the iteration.
  for (int i = 0; i < npts / 2; ++i) {
    const double m = 0.5 * (x[N - i] - x[i]);
    x[i] = -m;
    x[N - i] = m;
  }
  if (npts % 2 == 1) x[N / 2] = 0.0;
}

// Warp function of the warp-and-blend construction: the 1D map from
// equidistant to LGL points, evaluated through the equidistant Lagrange
// interpolant and divided by the blend (1 - r^2); zero at the end points.
static double warp_factor(int N, double rout) {
  std::vector<double> lgl, w;
  lgl_points(N + 1, lgl, w);
  std::vector<double> req(N + 1);
  for (int i = 0; i <= N; ++i) req[i] = -1.0 + 2.0 * i / N;
  req[0] = -1.0;
  req[N] = 1.0;
  // Veq(i, j) = P_j(req_i); Lagrange values l = Veq^{-T} P(rout).
  Mat VeqT(N + 1, N + 1);
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N; ++j) VeqT(j, i) = jacobi_p(req[i], 0.0, 0.0, j);
  Mat L = inverse(VeqT);
  double warp = 0.0;
  for (int i = 0; i <= N; ++i) {
    double li = 0.0;
    for (int j = 0; j <= N; ++j) li += L(i, j) * jacobi_p(rout, 0.0, 0.0, j);
    warp += li * (lgl[i] - req[i]);
  }
  if (std::fabs(rout) < 1.0 - 10.0 * std::numeric_limits<double>::epsilon())
    return warp / (1.0 - rout * rout);
  return 0.0;
}

void triangle_nodes(int N, std::vector<double>& r, std::vector<double>& s) {
  if (N < 1) throw std::invalid_argument("triangle_nodes: order must be >= 1");
  // Optimised blending parameters (Hesthaven & Warburton, Nodes2D, Table 6.1).
  static const double alpopt[15] = {0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9808, 1.0999, 1.2832,
                                    1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258};
  const double alpha = N < 16 ? alpopt[N - 1] : 5.0 / 3.0;
  const double sq3 = std::sqrt(3.0);
  r.clear();
  s.clear();
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N - i; ++j) {
      const double L1 = double(i) / N, L3 = double(j) / N, L2 = 1.0 - L1 - L3;
      double x = L3 - L2;
      double y = (2.0 * L1 - L2 - L3) / sq3;
      const double w1 = 4.0 * L2 * L3 * warp_factor(N, L3 - L2) * (1.0 + (alpha * L1) * (alpha * L1));
      const double w2 = 4.0 * L1 * L3 * warp_factor(N, L1 - L3) * (1.0 + (alpha * L2) * (alpha * L2));
      const double w3 = 4.0 * L1 * L2 * warp_factor(N, L2 - L1) * (1.0 + (alpha * L3) * (alpha * L3));
      x += w1 - 0.5 * w2 - 0.5 * w3;
      y += 0.5 * sq3 * (w2 - w3);
      // equilateral -> reference (r, s)
      const double l1 = (1.0 + sq3 * y) / 3.0;
      const double l2 = (2.0 - 3.0 * x - sq3 * y) / 6.0;
      const double l3 = (2.0 + 3.0 * x - sq3 * y) / 6.0;
      r.push_back(l3 - l2 - l1);
      s.push_back(l1 - l2 - l3);
    }
}

int face_node(int N, int face, int d) {
  const int Np = (N + 1) * (N + 2) / 2;
  switch (face) {
    case 0: return d;
    case 1: return N + d * N - d * (d - 1) / 2;
    case 2: return Np - 1 - d * (d + 3) / 2;
    default: throw std::invalid_argument("face_node: face must be 0..2");
  }
}

// Dubiner basis on the triangle in collapsed coordinates (a, b).
static void rs_to_ab(double r, double s, double& a, double& b) {
  a = (s == 1.0) ? -1.0 : 2.0 * (1.0 + r) / (1.0 - s) - 1.0;
  b = s;
}

static double simplex_p(double r, double s, int i, int j) {
  double a, b;
  rs_to_ab(r, s, a, b);
  return std::sqrt(2.0) * jacobi_p(a, 0, 0, i) * jacobi_p(b, 2.0 * i + 1.0, 0, j) * std::pow(1.0 - b, i);
}

static void grad_simplex_p(double r, double s, int i, int j, double& dr, double& ds) {
  double a, b;
  rs_to_ab(r, s, a, b);
  const double fa = jacobi_p(a, 0, 0, i), dfa = grad_jacobi_p(a, 0, 0, i);
  const double gb = jacobi_p(b, 2.0 * i + 1.0, 0, j), dgb = grad_jacobi_p(b, 2.0 * i + 1.0, 0, j);
  if (i == 0) {
    dr = std::sqrt(2.0) * dfa * gb;
    ds = std::sqrt(2.0) * (dfa * gb * 0.5 * (1.0 + a) + fa * dgb);
  } else {
    const double pb = std::pow(1.0 - b, i - 1);
    dr = std::sqrt(2.0) * 2.0 * dfa * gb * pb;
    ds = std::sqrt(2.0) * pb * (dfa * gb * (1.0 + a) + fa * dgb * (1.0 - b) - i * fa * gb);
  }
}

RefTriangle::RefTriangle(int order, OperatorForm form) : N(order) {
  if (order < 1) throw std::invalid_argument("RefTriangle: order must be >= 1");
  Np = (N + 1) * (N + 2) / 2;
  Nfp = N + 1;
  triangle_nodes(N, r, s);
  fmask.resize(3 * Nfp);
  for (int f = 0; f < 3; ++f)
    for (int d = 0; d < Nfp; ++d) fmask[f * Nfp + d] = face_node(N, f, d);

  V = Mat(Np, Np);
  Vr = Mat(Np, Np);
  Vs = Mat(Np, Np);
  int col = 0;
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N - i; ++j, ++col)
      for (int n = 0; n < Np; ++n) {
        V(n, col) = simplex_p(r[n], s[n], i, j);
        grad_simplex_p(r[n], s[n], i, j, Vr(n, col), Vs(n, col));
      }
  MassInv = matmul(V, transpose(V));
  Mass = inverse(MassInv);
  if (form == kWeakForm) {
    // M^{-1} S^T: Dr_w = V Vr^T M
    Dr = matmul(matmul(V, transpose(Vr)), Mass);
    Ds = matmul(matmul(V, transpose(Vs)), Mass);
  } else {
    const Mat Vinv = inverse(V);
    Dr = matmul(Vr, Vinv);
    Ds = matmul(Vs, Vinv);
  }
  // Face mass matrices scattered to face nodes, then LIFT = M^{-1} E.
  Mat E(Np, 3 * Nfp);
  for (int f = 0; f < 3; ++f) {
    Mat fv(Nfp, Nfp);
    for (int row = 0; row < Nfp; ++row) {
      const int n = fmask[f * Nfp + row];
      const double t = (f == 0) ? r[n] : s[n];
      for (int c = 0; c < Nfp; ++c) fv(row, c) = jacobi_p(t, 0, 0, c);
    }
    const Mat fm = inverse(matmul(fv, transpose(fv)));
    for (int row = 0; row < Nfp; ++row)
      for (int c = 0; c < Nfp; ++c) E(fmask[f * Nfp + row], f * Nfp + c) = fm(row, c);
  }
  LIFT = matmul(MassInv, E);
}

}  // namespace dgc_host
