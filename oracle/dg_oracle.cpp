// ORACLE -- test infrastructure only.
//
// A plain C++ restatement of the reference's serial CPU path (USE_CPU_ONLY,
// src/config.h:28-29) for the nodal-DG hot path. Only tests/, the smoke()
// check and the bench's cpu_baseline leg may load this library, and only as
// the checker: the product never links or calls it.
//
// Layout is the reference CPU layout: cell-major DOFs u[c * Np + j].
// Geometry is recomputed per call from the cell vertices the way the
// reference does (src/mapping_2d.h, src/basic_geom_2d.h), and the per-cell
// physical derivative matrices are formed explicitly (Dx = Dr*rx + Ds*sx)
// and applied with the row-major gemv order of src/dense_matrix.h:466-491.
//
// Pinning: tests/test_oracle.py checks every function here against the
// golden vectors written by the compiled reference (tests/golden/), and
// against closed forms (polynomial exactness, energy, symmetry).
//
// The cell loops carry OpenMP pragmas (per-cell work, no reductions, so a
// threaded build gives the same bits); liboracle.so is built without
// -fopenmp (serial, as the reference), liboracle_omp.so with it, for the
// all-core CPU baseline.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

namespace {

struct Cell {
  double x[3], y[3];
};

// src/alpha_optimized_nodes_triangle.h:69-102
int face_node(int N, int f, int d) {
  const int Np = (N + 1) * (N + 2) / 2;
  return f == 0 ? d : (f == 1 ? N + d * N - d * (d - 1) / 2 : Np - 1 - d * (d + 3) / 2);
}

Cell get_cell(const double* cx, const double* cy, int c) {
  Cell t;
  for (int v = 0; v < 3; ++v) {
    t.x[v] = cx[3 * c + v];
    t.y[v] = cy[3 * c + v];
  }
  return t;
}

// src/mapping_2d.h:120-129 (signed J)
double cell_J(const Cell& t) {
  return (1.0 / 4.0) * ((t.x[1] - t.x[0]) * (t.y[2] - t.y[0]) - (t.x[2] - t.x[0]) * (t.y[1] - t.y[0]));
}

// src/mapping_2d.h:131-158
double face_J(const Cell& t, int f) {
  const int a = f, b = (f + 1) % 3;
  const double dx = t.x[a] - t.x[b], dy = t.y[a] - t.y[b];
  return (1.0 / 2.0) * std::sqrt(dx * dx + dy * dy);
}

// src/basic_geom_2d.h:110-153
void outward_normal(const Cell& t, int e, double& nx, double& ny) {
  const int p = e, q = (e + 1) % 3, o = (e + 2) % 3;
  double x = t.x[q] - t.x[p], y = t.y[q] - t.y[p];
  const double x1 = t.x[o] - t.x[p], y1 = t.y[o] - t.y[p];
  if ((y * x1 - x * y1) > 0.0) {
    y = -y;
    x = -x;
  }
  const double l = std::sqrt(y * y + x * x);
  nx = y / l;
  ny = -x / l;
}

// src/mapping_2d.h:59-99 (rs -> xy)
void rs_to_xy(const Cell& t, double r, double s, double& x, double& y) {
  const double half = 1.0 / 2.0;
  const double l1 = half * (s + 1.0), l2 = -half * (r + s), l3 = half * (r + 1.0);
  x = l2 * t.x[0] + l3 * t.x[1] + l1 * t.x[2];
  y = l2 * t.y[0] + l3 * t.y[1] + l1 * t.y[2];
}

// src/simple_discretization_2d.h:84-102 with the Gauss-Jordan inverse of
// src/dense_matrix.h:505-566 applied to the 2x2 Jacobian of src/mapping_2d.h:101-118.
void inv_jacobian(const Cell& t, double out[4]) {
  double m[2][2] = {{0.5 * (t.x[1] - t.x[0]), 0.5 * (t.x[2] - t.x[0])},
                    {0.5 * (t.y[1] - t.y[0]), 0.5 * (t.y[2] - t.y[0])}};
  double inv[2][2] = {{1.0, 0.0}, {0.0, 1.0}};
  for (int i = 0; i < 2; ++i) {
    int piv = i;
    for (int ii = i + 1; ii < 2; ++ii)
      if (std::fabs(m[piv][i]) < std::fabs(m[ii][i])) piv = ii;
    if (piv != i)
      for (int j = 0; j < 2; ++j) {
        std::swap(m[i][j], m[piv][j]);
        std::swap(inv[i][j], inv[piv][j]);
      }
    double d = m[i][i];
    for (int j = 0; j < 2; ++j) {
      m[i][j] /= d;
      inv[i][j] /= d;
    }
    for (int ii = 0; ii < 2; ++ii)
      if (ii != i) {
        d = m[ii][i];
        for (int j = 0; j < 2; ++j) {
          m[ii][j] -= d * m[i][j];
          inv[ii][j] -= d * inv[i][j];
        }
      }
  }
  out[0] = inv[0][0];
  out[1] = inv[0][1];
  out[2] = inv[1][0];
  out[3] = inv[1][1];
}

// src/dense_matrix.h:466-491 (row-major): y = beta*y + sum_j alpha*m_ij*x_j
void gemv(const double* m, int rows, int cols, double alpha, const double* x, double beta, double* y) {
  for (int i = 0; i < rows; ++i) {
    double v = beta * y[i];
    for (int j = 0; j < cols; ++j) v += alpha * m[i * cols + j] * x[j];
    y[i] = v;
  }
}

// Dx = Dr * a + Ds * b (the reference's operator* then operator+)
void combine(const double* Dr, const double* Ds, int n, double a, double b, double* D) {
  for (int i = 0; i < n; ++i) D[i] = Dr[i] * a + Ds[i] * b;
}

struct Problem {
  int K, N, Np, Nfp;
  const double *Dr, *Ds, *L, *cx, *cy;
  const int *nb_cell, *nb_face;
  bool boundary(int c, int e) const { return nb_cell[3 * c + e] == c; }
};

}  // namespace

extern "C" {

// Maxwell TM RHS: examples/maxwell_2d/maxwell_2d.h:179-311.
// in/out: 3 fields (Hx, Hy, Ez) each K*Np cell-major, at in + f*K*Np.
void oracle_maxwell_rhs(int K, int N, const double* Dr, const double* Ds, const double* L,
                        const double* cx, const double* cy, const int* nb_cell, const int* nb_face,
                        const double* in, double t, double* out) {
  (void)t;
  const int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1;
  const size_t n = size_t(K) * Np;
  const double *inHx = in, *inHy = in + n, *inEz = in + 2 * n;
  double *outHx = out, *outHy = out + n, *outEz = out + 2 * n;
  std::vector<double> nfHx(size_t(K) * 3 * Nfp), nfHy(nfHx.size()), nfEz(nfHx.size());
  // numerical fluxes (maxwell_2d.h:179-256)
#pragma omp parallel for schedule(static)
  for (int c = 0; c < K; ++c) {
    const Cell cell = get_cell(cx, cy, c);
    for (int e = 0; e < 3; ++e) {
      const bool isB = nb_cell[3 * c + e] == c;
      const int nbc = nb_cell[3 * c + e], nbf = nb_face[3 * c + e];
      for (int d = 0; d < Nfp; ++d) {
        const int a = c * Np + face_node(N, e, d);
        const double aHx = inHx[a], aHy = inHy[a], aEz = inEz[a];
        double bHx, bHy, bEz;
        if (isB) {
          bHx = aHx;
          bHy = aHy;
          bEz = -aEz;
        } else {
          const int b = nbc * Np + face_node(N, nbf, Nfp - d - 1);
          bHx = inHx[b];
          bHy = inHy[b];
          bEz = inEz[b];
        }
        const int o = c * 3 * Nfp + e * Nfp + d;
        double nx, ny;
        outward_normal(cell, e, nx, ny);
        const double dltHx = aHx - bHx, dltHy = aHy - bHy, sumEz = aEz + bEz;
        nfHx[o] = 0.5 * ny * (sumEz + ny * dltHx - nx * dltHy);
        nfHy[o] = -0.5 * nx * (sumEz + ny * dltHx - nx * dltHy);
        nfEz[o] = 0.5 * (ny * (aHx + bHx) - nx * (aHy + bHy) + aEz - bEz);
      }
    }
  }
#pragma omp parallel
  {
  std::vector<double> Dx(Np * Np), Dy(Np * Np), fHx(3 * Nfp), fHy(3 * Nfp), fEz(3 * Nfp);
#pragma omp for schedule(static)
  for (int c = 0; c < K; ++c) {
    const Cell cell = get_cell(cx, cy, c);
    double ij[4];
    inv_jacobian(cell, ij);
    combine(Dr, Ds, Np * Np, ij[0], ij[2], Dx.data());
    combine(Dr, Ds, Np * Np, ij[1], ij[3], Dy.data());
    const size_t o = size_t(c) * Np;
    gemv(Dy.data(), Np, Np, 1.0, inEz + o, 0.0, outHx + o);
    gemv(Dx.data(), Np, Np, -1.0, inEz + o, 0.0, outHy + o);
    gemv(Dx.data(), Np, Np, -1.0, inHy + o, 0.0, outEz + o);
    gemv(Dy.data(), Np, Np, 1.0, inHx + o, 1.0, outEz + o);
    for (int e = 0; e < 3; ++e) {
      const double fj = face_J(cell, e);
      for (int i = 0; i < Nfp; ++i) {
        const int ei = e * Nfp + i;
        fHx[ei] = nfHx[size_t(c) * 3 * Nfp + ei] * fj;
        fHy[ei] = nfHy[size_t(c) * 3 * Nfp + ei] * fj;
        fEz[ei] = nfEz[size_t(c) * 3 * Nfp + ei] * fj;
      }
    }
    const double J = cell_J(cell);
    gemv(L, Np, 3 * Nfp, -1.0 / J, fHx.data(), 1.0, outHx + o);
    gemv(L, Np, 3 * Nfp, -1.0 / J, fHy.data(), 1.0, outHy + o);
    gemv(L, Np, 3 * Nfp, -1.0 / J, fEz.data(), 1.0, outEz + o);
  }
  }
}

// Advection RHS: examples/advection_2d/advection_2d.h:156-258, generalised
// from a = (1, 1) to a = (ax, ay) (a = (1,1) reproduces the reference
// bit-for-bit: the extra products are by 1.0). bc 0: analytic inflow
// sin(pi(x+1-t)) sin(pi(y+1-t)); bc 1: zero inflow. r, s: reference nodes.
void oracle_advection_rhs(int K, int N, const double* Dr, const double* Ds, const double* L,
                          const double* cx, const double* cy, const int* nb_cell,
                          const int* nb_face, const double* r, const double* s, double ax,
                          double ay, int bc, const double* in, double t, double* out) {
  const int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1;
  std::vector<double> nf(size_t(K) * 3 * Nfp);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < K; ++c) {
    const Cell cell = get_cell(cx, cy, c);
    for (int e = 0; e < 3; ++e) {
      const bool isB = nb_cell[3 * c + e] == c;
      const int nbc = nb_cell[3 * c + e], nbf = nb_face[3 * c + e];
      for (int d = 0; d < Nfp; ++d) {
        const int fn = face_node(N, e, d);
        const double aU = in[size_t(c) * Np + fn];
        double bU;
        if (isB) {
          if (bc == 0) {
            double x, y;
            rs_to_xy(cell, r[fn], s[fn], x, y);
            bU = std::sin((x + 1.0 - t) * M_PI) * std::sin((y + 1.0 - t) * M_PI);
          } else {
            bU = 0.0;
          }
        } else {
          bU = in[size_t(nbc) * Np + face_node(N, nbf, Nfp - d - 1)];
        }
        double nx, ny;
        outward_normal(cell, e, nx, ny);
        const double an = ax * nx + ay * ny;
        const double U = an >= 0.0 ? aU : bU;
        nf[size_t(c) * 3 * Nfp + e * Nfp + d] = U * an;
      }
    }
  }
#pragma omp parallel
  {
  std::vector<double> Dx(Np * Np), Dy(Np * Np), mFl(3 * Nfp);
#pragma omp for schedule(static)
  for (int c = 0; c < K; ++c) {
    const Cell cell = get_cell(cx, cy, c);
    double ij[4];
    inv_jacobian(cell, ij);
    combine(Dr, Ds, Np * Np, ij[0], ij[2], Dx.data());
    combine(Dr, Ds, Np * Np, ij[1], ij[3], Dy.data());
    const size_t o = size_t(c) * Np;
    gemv(Dx.data(), Np, Np, ax, in + o, 0.0, out + o);
    gemv(Dy.data(), Np, Np, ay, in + o, 1.0, out + o);
    for (int e = 0; e < 3; ++e) {
      const double fj = face_J(cell, e);
      for (int i = 0; i < Nfp; ++i) mFl[e * Nfp + i] = nf[size_t(c) * 3 * Nfp + e * Nfp + i] * fj;
    }
    gemv(L, Np, 3 * Nfp, -1.0 / cell_J(cell), mFl.data(), 1.0, out + o);
  }
  }
}

// Penalty factor per face. tau_mode 0: the reference's tau = 1
// (src/laplace_2d.h:227-245); tau_mode 1: tau = (N+1)(N+2)/h with
// h = min over the two sides of 2J/face_J (the cell height to that face).
static double face_tau(int tau_mode, int N, const double* cx, const double* cy, const int* nb_cell,
                       const int* nb_face, int c, int e) {
  if (tau_mode == 0) return 1.0;
  const Cell a = get_cell(cx, cy, c);
  double h = 2.0 * cell_J(a) / face_J(a, e);
  const int nc = nb_cell[3 * c + e];
  if (nc != c) {
    const Cell b = get_cell(cx, cy, nc);
    const double hb = 2.0 * cell_J(b) / face_J(b, nb_face[3 * c + e]);
    if (hb < h) h = hb;
  }
  return (N + 1.0) * (N + 2.0) / h;
}

// IPDG Laplacian with homogeneous Dirichlet data, in place:
// src/laplace_2d.h:121-251 with the BC of examples/poisson_2d/poisson_2d.h:73-79
// (every boundary is Dirichlet; the inhomogeneous part is zero for g = 0).
// Strong-form Dr, Ds. u: K*Np cell-major, overwritten by Delta_h u.
void oracle_laplace_apply_data(int K, int N, const double* Dr, const double* Ds, const double* L,
                               const double* cx, const double* cy, const int* nb_cell, const int* nb_face,
                               int tau_mode, int neumann, const double* gd, const double* hd, double* u);
static std::vector<int> boundary_offsets(int K, int Nfp, const int* nb_cell);
void oracle_laplace_apply_bc(int K, int N, const double* Dr, const double* Ds, const double* L,
                             const double* cx, const double* cy, const int* nb_cell,
                             const int* nb_face, int tau_mode, int neumann, double* u);

void oracle_laplace_apply(int K, int N, const double* Dr, const double* Ds, const double* L,
                          const double* cx, const double* cy, const int* nb_cell,
                          const int* nb_face, int tau_mode, double* u) {
  oracle_laplace_apply_bc(K, N, Dr, Ds, L, cx, cy, nb_cell, nb_face, tau_mode, 0, u);
}

// laplace_2d::apply_homogeneous (src/laplace_2d.h:130-251) with one boundary
// kind for every boundary node: neumann 0 is bc.is_dirichlet() == true
// (du = 2u, dqN = 0), neumann 1 is false (du = 0, dqN = 2 q.n).
void oracle_laplace_apply_bc(int K, int N, const double* Dr, const double* Ds, const double* L,
                             const double* cx, const double* cy, const int* nb_cell,
                             const int* nb_face, int tau_mode, int neumann, double* u) {
  oracle_laplace_apply_data(K, N, Dr, Ds, L, cx, cy, nb_cell, nb_face, tau_mode, neumann, nullptr, nullptr, u);
}

// The same two passes with boundary data (gd: the exterior value at interior
// value 0 on Dirichlet faces, u+ = gd - u-; hd: the exterior normal
// derivative at interior 0 on Neumann faces, q+.n = hd - q-.n; per boundary
// node, NULL = 0): the whole BR1 operator with inhomogeneous data, whose
// affine part also reaches the interior faces of boundary cells (which
// laplace_2d::apply_inhomogeneous_as_rhs leaves out).
void oracle_laplace_apply_data(int K, int N, const double* Dr, const double* Ds, const double* L,
                               const double* cx, const double* cy, const int* nb_cell, const int* nb_face,
                               int tau_mode, int neumann, const double* gd, const double* hd, double* u) {
  const int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1;
  const std::vector<int> boff = boundary_offsets(K, Nfp, nb_cell);
  std::vector<double> qx(size_t(K) * Np), qy(size_t(K) * Np), du(size_t(K) * 3 * Nfp);
  std::vector<double> Dx(Np * Np), Dy(Np * Np), tlx(3 * Nfp), tly(3 * Nfp);
  // pass 1: q = grad u with central flux (laplace_2d.h:145-197)
  for (int c = 0; c < K; ++c) {
    const Cell cell = get_cell(cx, cy, c);
    double ij[4];
    inv_jacobian(cell, ij);
    combine(Dr, Ds, Np * Np, ij[0], ij[2], Dx.data());
    combine(Dr, Ds, Np * Np, ij[1], ij[3], Dy.data());
    const size_t o = size_t(c) * Np;
    gemv(Dx.data(), Np, Np, 1.0, u + o, 0.0, qx.data() + o);
    gemv(Dy.data(), Np, Np, 1.0, u + o, 0.0, qy.data() + o);
    for (int e = 0; e < 3; ++e) {
      const bool isB = nb_cell[3 * c + e] == c;
      const int nbc = nb_cell[3 * c + e], nbf = nb_face[3 * c + e];
      double nx, ny;
      outward_normal(cell, e, nx, ny);
      const double fj = face_J(cell, e);
      for (int i = 0; i < Nfp; ++i) {
        double d;
        if (isB)
          d = neumann ? 0.0 : 2.0 * u[o + face_node(N, e, i)] - (gd ? gd[boff[3 * c + e] + i] : 0.0);
        else
          d = u[o + face_node(N, e, i)] - u[size_t(nbc) * Np + face_node(N, nbf, Nfp - i - 1)];
        du[size_t(c) * 3 * Nfp + e * Nfp + i] = d;
        tlx[e * Nfp + i] = nx * d * fj / 2.0;
        tly[e * Nfp + i] = ny * d * fj / 2.0;
      }
    }
    const double J = cell_J(cell);
    gemv(L, Np, 3 * Nfp, -1.0 / J, tlx.data(), 1.0, qx.data() + o);
    gemv(L, Np, 3 * Nfp, -1.0 / J, tly.data(), 1.0, qy.data() + o);
  }
  // pass 2: Delta u = div q with q* = {q} - tau [u] (laplace_2d.h:199-250)
  for (int c = 0; c < K; ++c) {
    const Cell cell = get_cell(cx, cy, c);
    double ij[4];
    inv_jacobian(cell, ij);
    combine(Dr, Ds, Np * Np, ij[0], ij[2], Dx.data());
    combine(Dr, Ds, Np * Np, ij[1], ij[3], Dy.data());
    const size_t o = size_t(c) * Np;
    gemv(Dx.data(), Np, Np, 1.0, qx.data() + o, 0.0, u + o);
    gemv(Dy.data(), Np, Np, 1.0, qy.data() + o, 1.0, u + o);
    for (int e = 0; e < 3; ++e) {
      const bool isB = nb_cell[3 * c + e] == c;
      const int nbc = nb_cell[3 * c + e], nbf = nb_face[3 * c + e];
      double nx, ny;
      outward_normal(cell, e, nx, ny);
      const double fj = face_J(cell, e);
      const double tau = face_tau(tau_mode, N, cx, cy, nb_cell, nb_face, c, e);
      for (int i = 0; i < Nfp; ++i) {
        const int a = face_node(N, e, i);
        double dqN;
        if (isB) {
          dqN = neumann ? 2.0 * (qx[o + a] * nx + qy[o + a] * ny) - (hd ? hd[boff[3 * c + e] + i] : 0.0) : 0.0;
        } else {
          const size_t b = size_t(nbc) * Np + face_node(N, nbf, Nfp - i - 1);
          dqN = (qx[o + a] - qx[b]) * nx + (qy[o + a] - qy[b]) * ny;
        }
        const double d = du[size_t(c) * 3 * Nfp + e * Nfp + i];
        tlx[e * Nfp + i] = (dqN + 2.0 * (tau_mode == 0 ? d : tau * d)) * fj / 2.0;
      }
    }
    gemv(L, Np, 3 * Nfp, -1.0 / cell_J(cell), tlx.data(), 1.0, u + o);
  }
}

static std::vector<int> boundary_offsets(int K, int Nfp, const int* nb_cell);

// laplace_2d::apply_inhomogeneous_as_rhs (src/laplace_2d.h:253-330): inout +=
// the boundary part for data gu (exterior_val(x, y, 0) on Dirichlet faces) or
// hq (exterior_grad_n(x, y, 0) on Neumann faces) per boundary node (NULL = 0),
// one boundary kind for every face as oracle_laplace_apply_bc.
void oracle_laplace_inhomogeneous(int K, int N, const double* Dr, const double* Ds, const double* L,
                                  const double* cx, const double* cy, const int* nb_cell, const int* nb_face,
                                  int tau_mode, int neumann, const double* gu, const double* hq, double* inout) {
  const int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1;
  const std::vector<int> boff = boundary_offsets(K, Nfp, nb_cell);
  std::vector<double> qx(Np), qy(Np), du(3 * Nfp), Dx(Np * Np), Dy(Np * Np), tlx(3 * Nfp), tly(3 * Nfp);
  for (int c = 0; c < K; ++c) {
    if (nb_cell[3 * c] != c && nb_cell[3 * c + 1] != c && nb_cell[3 * c + 2] != c) continue;
    const Cell cell = get_cell(cx, cy, c);
    std::fill(tlx.begin(), tlx.end(), 0.0);
    std::fill(tly.begin(), tly.end(), 0.0);
    std::fill(du.begin(), du.end(), 0.0);
    for (int e = 0; e < 3; ++e) {
      if (nb_cell[3 * c + e] != c) continue;
      double nx, ny;
      outward_normal(cell, e, nx, ny);
      const double fj = face_J(cell, e);
      for (int i = 0; i < Nfp; ++i) {
        const double g = (!neumann && gu) ? gu[boff[3 * c + e] + i] : 0.0;
        du[e * Nfp + i] = g;
        tlx[e * Nfp + i] = nx * g * fj / 2.0;
        tly[e * Nfp + i] = ny * g * fj / 2.0;
      }
    }
    const double J = cell_J(cell);
    gemv(L, Np, 3 * Nfp, -1.0 / J, tlx.data(), 0.0, qx.data());
    gemv(L, Np, 3 * Nfp, -1.0 / J, tly.data(), 0.0, qy.data());
    double ij[4];
    inv_jacobian(cell, ij);
    combine(Dr, Ds, Np * Np, ij[0], ij[2], Dx.data());
    combine(Dr, Ds, Np * Np, ij[1], ij[3], Dy.data());
    const size_t o = size_t(c) * Np;
    gemv(Dx.data(), Np, Np, 1.0, qx.data(), 1.0, inout + o);
    gemv(Dy.data(), Np, Np, 1.0, qy.data(), 1.0, inout + o);
    std::fill(tlx.begin(), tlx.end(), 0.0);
    for (int e = 0; e < 3; ++e) {
      if (nb_cell[3 * c + e] != c) continue;
      const double fj = face_J(cell, e);
      const double tau = face_tau(tau_mode, N, cx, cy, nb_cell, nb_face, c, e);
      for (int i = 0; i < Nfp; ++i) {
        const double h = (neumann && hq) ? hq[boff[3 * c + e] + i] : 0.0;
        const double d = du[e * Nfp + i];
        tlx[e * Nfp + i] = (h + 2.0 * (tau_mode == 0 ? d : tau * d)) * fj / 2.0;
      }
    }
    gemv(L, Np, 3 * Nfp, -1.0 / J, tlx.data(), 1.0, inout + o);
  }
}

// Boundary-node offsets of the faces of boundary half-edges, counted cell by
// cell, face by face (the order the reference calls its BC objects in); -1 for
// interior faces.
static std::vector<int> boundary_offsets(int K, int Nfp, const int* nb_cell) {
  std::vector<int> off(size_t(3) * K, -1);
  int b = 0;
  for (int c = 0; c < K; ++c)
    for (int e = 0; e < 3; ++e)
      if (nb_cell[3 * c + e] == c) {
        off[3 * c + e] = b;
        b += Nfp;
      }
  return off;
}

// Weak gradient with central flux: src/grad_2d.h:95-187. The BC object is given
// in its affine form: exterior value g[b] + beta * interior at boundary node b
// (g == NULL: 0). Weak-form Dr, Ds. in: K*Np cell-major; out: ox, oy.
void oracle_grad(int K, int N, const double* Dr, const double* Ds, const double* L, const double* cx,
                 const double* cy, const int* nb_cell, const int* nb_face, const double* g, double beta,
                 const double* in, double* ox, double* oy) {
  const int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1;
  const std::vector<int> boff = boundary_offsets(K, Nfp, nb_cell);
  std::vector<double> nf(size_t(K) * 3 * Nfp);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < K; ++c)
    for (int e = 0; e < 3; ++e)
      for (int d = 0; d < Nfp; ++d) {
        const double aU = in[size_t(c) * Np + face_node(N, e, d)];
        double bU;
        if (nb_cell[3 * c + e] == c)
          bU = (g ? g[boff[3 * c + e] + d] : 0.0) + beta * aU;
        else
          bU = in[size_t(nb_cell[3 * c + e]) * Np + face_node(N, nb_face[3 * c + e], Nfp - d - 1)];
        nf[size_t(c) * 3 * Nfp + e * Nfp + d] = (aU + bU) / 2.0;
      }
#pragma omp parallel
  {
  std::vector<double> Dx(Np * Np), Dy(Np * Np), fx(3 * Nfp), fy(3 * Nfp);
#pragma omp for schedule(static)
  for (int c = 0; c < K; ++c) {
    const Cell cell = get_cell(cx, cy, c);
    double ij[4];
    inv_jacobian(cell, ij);
    combine(Dr, Ds, Np * Np, ij[0], ij[2], Dx.data());
    combine(Dr, Ds, Np * Np, ij[1], ij[3], Dy.data());
    const size_t o = size_t(c) * Np;
    gemv(Dx.data(), Np, Np, -1.0, in + o, 0.0, ox + o);
    gemv(Dy.data(), Np, Np, -1.0, in + o, 0.0, oy + o);
    for (int e = 0; e < 3; ++e) {
      const double fj = face_J(cell, e);
      double nx, ny;
      outward_normal(cell, e, nx, ny);
      for (int i = 0; i < Nfp; ++i) {
        fx[e * Nfp + i] = nf[size_t(c) * 3 * Nfp + e * Nfp + i] * nx * fj;
        fy[e * Nfp + i] = nf[size_t(c) * 3 * Nfp + e * Nfp + i] * ny * fj;
      }
    }
    gemv(L, Np, 3 * Nfp, 1.0 / cell_J(cell), fx.data(), 1.0, ox + o);
    gemv(L, Np, 3 * Nfp, 1.0 / cell_J(cell), fy.data(), 1.0, oy + o);
  }
  }
}

// Weak divergence with central flux: src/div_2d.h:95-188, BC as oracle_grad
// with boundary data g0, g1 for the two components. out: K*Np.
void oracle_div(int K, int N, const double* Dr, const double* Ds, const double* L, const double* cx,
                const double* cy, const int* nb_cell, const int* nb_face, const double* g0, const double* g1,
                double beta, const double* u0, const double* u1, double* out) {
  const int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1;
  const std::vector<int> boff = boundary_offsets(K, Nfp, nb_cell);
  std::vector<double> nf(size_t(K) * 3 * Nfp);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < K; ++c) {
    const Cell cell = get_cell(cx, cy, c);
    for (int e = 0; e < 3; ++e) {
      double nx, ny;
      outward_normal(cell, e, nx, ny);
      for (int d = 0; d < Nfp; ++d) {
        const size_t ia = size_t(c) * Np + face_node(N, e, d);
        const double a0 = u0[ia], a1 = u1[ia];
        double b0, b1;
        if (nb_cell[3 * c + e] == c) {
          const int b = boff[3 * c + e] + d;
          b0 = (g0 ? g0[b] : 0.0) + beta * a0;
          b1 = (g1 ? g1[b] : 0.0) + beta * a1;
        } else {
          const size_t ib = size_t(nb_cell[3 * c + e]) * Np + face_node(N, nb_face[3 * c + e], Nfp - d - 1);
          b0 = u0[ib];
          b1 = u1[ib];
        }
        nf[size_t(c) * 3 * Nfp + e * Nfp + d] = ((a0 + b0) * nx + (a1 + b1) * ny) / 2.0;
      }
    }
  }
#pragma omp parallel
  {
  std::vector<double> Dx(Np * Np), Dy(Np * Np), fl(3 * Nfp);
#pragma omp for schedule(static)
  for (int c = 0; c < K; ++c) {
    const Cell cell = get_cell(cx, cy, c);
    double ij[4];
    inv_jacobian(cell, ij);
    combine(Dr, Ds, Np * Np, ij[0], ij[2], Dx.data());
    combine(Dr, Ds, Np * Np, ij[1], ij[3], Dy.data());
    const size_t o = size_t(c) * Np;
    gemv(Dx.data(), Np, Np, -1.0, u0 + o, 0.0, out + o);
    gemv(Dy.data(), Np, Np, -1.0, u1 + o, 1.0, out + o);
    for (int e = 0; e < 3; ++e) {
      const double fj = face_J(cell, e);
      for (int i = 0; i < Nfp; ++i) fl[e * Nfp + i] = nf[size_t(c) * 3 * Nfp + e * Nfp + i] * fj;
    }
    gemv(L, Np, 3 * Nfp, 1.0 / cell_J(cell), fl.data(), 1.0, out + o);
  }
  }
}

// 1D advection right-hand side: examples/advection_1d/advection_1d.h:155-197
// (u_t + a u_x = 0, a = 2 pi, inflow -sin(2 pi t), outflow copy, upwind flux).
// D, L: the strong-form m_D, m_L ((N+1)^2, row-major); in/out cell-major.
void oracle_adv1d_rhs(int K, int N, const double* D, const double* L, const double* in, double t, double* out) {
  const int np = N + 1;
  const double a = 2.0 * M_PI;
  auto flux = [a](double l, double r) { return a * 0.5 * (l + r) + a * 0.5 * (1.0 - 0.0) * (l - r); };
  std::vector<double> F(K + 1), iv(np), ov(np);
  for (int i = 0; i <= K; ++i) {
    const double l = i > 0 ? in[i * np - 1] : -std::sin(2.0 * M_PI * t);
    const double r = i < K ? in[i * np] : in[i * np - 1];
    F[i] = flux(l, r);
  }
  for (int c = 0; c < K; ++c) {
    gemv(D, np, np, -a, in + c * np, 0.0, out + c * np);
    std::fill(iv.begin(), iv.end(), 0.0);
    std::fill(ov.begin(), ov.end(), 0.0);
    iv[0] = F[c] - a * in[c * np];
    iv[np - 1] = a * in[(c + 1) * np - 1] - F[c + 1];
    gemv(L, np, np, 1.0, iv.data(), 1.0, ov.data());
    for (int j = 0; j < np; ++j) out[c * np + j] += ov[j];
  }
}

// nsteps of the RK4 of src/explicit_runge_kutta.h:33-51 (host axpy_n,
// src/axpy.h:112-122) on the 1D advection operator, in place from t0.
void oracle_adv1d_rk4(int K, int N, const double* D, const double* L, double* u, double t0, double dt, int nsteps) {
  const size_t n = size_t(K) * (N + 1);
  std::vector<double> w0(n), w1(n), w2(n), w3(n), w4(n);
  auto axpy = [n](double a, const double* x, const double* y, double* o) {
    for (size_t i = 0; i < n; ++i) o[i] = a * x[i] + y[i];
  };
  double t = t0;
  for (int s = 0; s < nsteps; ++s) {
    oracle_adv1d_rhs(K, N, D, L, u, t, w1.data());
    axpy(0.5 * dt, w1.data(), u, w0.data());
    oracle_adv1d_rhs(K, N, D, L, w0.data(), t + 0.5 * dt, w2.data());
    axpy(0.5 * dt, w2.data(), u, w0.data());
    oracle_adv1d_rhs(K, N, D, L, w0.data(), t + 0.5 * dt, w3.data());
    axpy(dt, w3.data(), u, w0.data());
    oracle_adv1d_rhs(K, N, D, L, w0.data(), t + dt, w4.data());
    axpy(2.0, w2.data(), w1.data(), w0.data());
    axpy(2.0, w3.data(), w4.data(), w1.data());
    axpy(dt / 6.0, w0.data(), u, w2.data());
    axpy(dt / 6.0, w1.data(), w2.data(), u);
    t += dt;
  }
}

// Classical RK4 of src/explicit_runge_kutta.h:33-51 with the host axpy_n of
// src/axpy.h:112-122, for either physics (phys 0: Maxwell, 1: advection).
// u: F*K*Np cell-major, advanced in place over `nsteps` steps of dt from t0;
// if t_end > t0 the loop instead runs `while (t < t_end)` with a truncated
// last step (examples/maxwell_2d/main.cpp:199-209). Returns steps taken.
int oracle_rk4(int phys, int K, int N, const double* Dr, const double* Ds, const double* L,
               const double* cx, const double* cy, const int* nb_cell, const int* nb_face,
               const double* r, const double* s, double ax, double ay, int bc, double* u, double t0,
               double dt, int nsteps, double t_end) {
  const int Np = (N + 1) * (N + 2) / 2;
  const int F = phys == 0 ? 3 : 1;
  const size_t n = size_t(F) * K * Np;
  std::vector<double> w0(n), w1(n), w2(n), w3(n), w4(n);
  auto op = [&](const double* in, double t, double* out) {
    if (phys == 0)
      oracle_maxwell_rhs(K, N, Dr, Ds, L, cx, cy, nb_cell, nb_face, in, t, out);
    else
      oracle_advection_rhs(K, N, Dr, Ds, L, cx, cy, nb_cell, nb_face, r, s, ax, ay, bc, in, t, out);
  };
  auto axpy = [&](double a, const double* x, const double* y, double* o) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) o[i] = a * x[i] + y[i];
  };
  auto step = [&](double t, double h) {
    op(u, t, w1.data());
    axpy(0.5 * h, w1.data(), u, w0.data());
    op(w0.data(), t + 0.5 * h, w2.data());
    axpy(0.5 * h, w2.data(), u, w0.data());
    op(w0.data(), t + 0.5 * h, w3.data());
    axpy(h, w3.data(), u, w0.data());
    op(w0.data(), t + h, w4.data());
    axpy(2.0, w2.data(), w1.data(), w0.data());
    axpy(2.0, w3.data(), w4.data(), w1.data());
    axpy(h / 6.0, w0.data(), u, w2.data());
    axpy(h / 6.0, w1.data(), w2.data(), u);
  };
  double t = t0;
  int steps = 0;
  if (t_end > t0) {
    while (t < t_end) {
      if (t + dt > t_end) dt = t_end - t;
      step(t, dt);
      t += dt;
      ++steps;
    }
  } else {
    for (int i = 0; i < nsteps; ++i) {
      step(t, dt);
      t += dt;
      ++steps;
    }
  }
  return steps;
}

}  // extern "C"
