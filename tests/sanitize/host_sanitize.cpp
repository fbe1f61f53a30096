// Host-code check under AddressSanitizer and UndefinedBehaviorSanitizer
// (test infrastructure; built by `make -C dg-on-cuda_amd sanitize`, run by
// tests/test_sanitizers.py). It drives the library's host set-up through its
// C ABI (reference element, structured / periodic / gmsh meshes, geometry,
// DOF coordinates, mesh and plot-file writers, error paths) and every
// function of the oracle on small meshes; any invalid access, leak or
// undefined operation aborts the run with the sanitizer's report.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/dgc_hip.h"

// The library's error slot (defined with the device code in the full build).
static std::string g_error;
int dgc_set_error(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

extern "C" {
void oracle_maxwell_rhs(int K, int N, const double* Dr, const double* Ds, const double* L, const double* cx,
                        const double* cy, const int* nb_cell, const int* nb_face, const double* in, double t,
                        double* out);
void oracle_advection_rhs(int K, int N, const double* Dr, const double* Ds, const double* L, const double* cx,
                          const double* cy, const int* nb_cell, const int* nb_face, const double* r, const double* s,
                          double ax, double ay, int bc, const double* in, double t, double* out);
void oracle_laplace_apply(int K, int N, const double* Dr, const double* Ds, const double* L, const double* cx,
                          const double* cy, const int* nb_cell, const int* nb_face, int tau_mode, double* u);
void oracle_laplace_apply_bc(int K, int N, const double* Dr, const double* Ds, const double* L, const double* cx,
                             const double* cy, const int* nb_cell, const int* nb_face, int tau_mode, int neumann,
                             double* u);
void oracle_laplace_apply_data(int K, int N, const double* Dr, const double* Ds, const double* L, const double* cx,
                               const double* cy, const int* nb_cell, const int* nb_face, int tau_mode, int neumann,
                               const double* gd, const double* hd, double* u);
void oracle_laplace_inhomogeneous(int K, int N, const double* Dr, const double* Ds, const double* L,
                                  const double* cx, const double* cy, const int* nb_cell, const int* nb_face,
                                  int tau_mode, int neumann, const double* gu, const double* hq, double* inout);
void oracle_grad(int K, int N, const double* Dr, const double* Ds, const double* L, const double* cx,
                 const double* cy, const int* nb_cell, const int* nb_face, const double* g, double beta,
                 const double* in, double* ox, double* oy);
void oracle_div(int K, int N, const double* Dr, const double* Ds, const double* L, const double* cx,
                const double* cy, const int* nb_cell, const int* nb_face, const double* g0, const double* g1,
                double beta, const double* u0, const double* u1, double* out);
void oracle_adv1d_rhs(int K, int N, const double* D, const double* L, const double* in, double t, double* out);
void oracle_adv1d_rk4(int K, int N, const double* D, const double* L, double* u, double t0, double dt, int nsteps);
int oracle_rk4(int phys, int K, int N, const double* Dr, const double* Ds, const double* L, const double* cx,
               const double* cy, const int* nb_cell, const int* nb_face, const double* r, const double* s, double ax,
               double ay, int bc, double* u, double t0, double dt, int nsteps, double t_end);
}

static int failures = 0;
#define CHECK(cond)                                                             \
  do {                                                                          \
    if (!(cond)) {                                                              \
      std::fprintf(stderr, "check failed at line %d: %s (%s)\n", __LINE__, #cond, g_error.c_str()); \
      ++failures;                                                               \
    }                                                                           \
  } while (0)

struct Ops {
  int N, Np, Nfp;
  std::vector<double> r, s, Dr, Ds, L, M;
  std::vector<int> fmask;
  Ops(int order, int form) : N(order), Np((order + 1) * (order + 2) / 2), Nfp(order + 1) {
    r.resize(Np);
    s.resize(Np);
    Dr.resize(Np * Np);
    Ds.resize(Np * Np);
    M.resize(Np * Np);
    L.resize(Np * 3 * Nfp);
    fmask.resize(3 * Nfp);
    CHECK(dgc_ref_operators(order, form, r.data(), s.data(), fmask.data(), Dr.data(), Ds.data(), L.data(),
                            M.data()) == DGC_OK);
  }
};

struct Cells {
  int K = 0, Nv = 0;
  double min_edge = 0;
  std::vector<double> cx, cy;
  std::vector<int> etov, nb_cell, nb_face;
  explicit Cells(dgc_mesh_t m) {
    CHECK(dgc_mesh_info(m, &K, &Nv, &min_edge) == DGC_OK);
    cx.resize(3 * K);
    cy.resize(3 * K);
    etov.resize(3 * K);
    nb_cell.resize(3 * K);
    nb_face.resize(3 * K);
    CHECK(dgc_mesh_cells(m, cx.data(), cy.data(), etov.data(), nb_cell.data(), nb_face.data()) == DGC_OK);
  }
};

static void exercise_mesh(dgc_mesh_t m, int order) {
  Cells c(m);
  const Ops w(order, 0), st(order, 1);
  const int K = c.K, Np = w.Np;
  const size_t n = size_t(K) * Np;
  // geometry: boundary-node count first, then the arrays
  std::vector<double> inv(4 * K), J(K), fJ(3 * K), nx(3 * K), ny(3 * K);
  std::vector<int> ic(3 * K), ifc(3 * K);
  int nb = 0;
  CHECK(dgc_mesh_geometry(m, order, inv.data(), J.data(), fJ.data(), nx.data(), ny.data(), ic.data(), ifc.data(),
                          &nb, nullptr, nullptr) == DGC_OK);
  std::vector<double> bx(nb + 1), by(nb + 1);
  CHECK(dgc_mesh_geometry(m, order, inv.data(), J.data(), fJ.data(), nx.data(), ny.data(), ic.data(), ifc.data(),
                          &nb, bx.data(), by.data()) == DGC_OK);
  std::vector<double> x(n), y(n);
  CHECK(dgc_mesh_dof_coords(m, order, 0, x.data(), y.data()) == DGC_OK);
  CHECK(dgc_mesh_dof_coords(m, order, 1, x.data(), y.data()) == DGC_OK);
  CHECK(dgc_mesh_dof_coords(m, order, 0, x.data(), y.data()) == DGC_OK);
  std::vector<double> u(3 * n), out(3 * n), v(n), a(n), b(n), g(nb + 1, 0.1);
  for (size_t i = 0; i < 3 * n; ++i) u[i] = std::sin(0.37 * double(i));
  for (size_t i = 0; i < n; ++i) v[i] = std::cos(1.3 * x[i]) * std::sin(0.7 * y[i]);
  const double *cx = c.cx.data(), *cy = c.cy.data();
  const int *nbc = c.nb_cell.data(), *nbf = c.nb_face.data();
  oracle_maxwell_rhs(K, order, w.Dr.data(), w.Ds.data(), w.L.data(), cx, cy, nbc, nbf, u.data(), 0.1, out.data());
  for (int bc = 0; bc < 2; ++bc)
    oracle_advection_rhs(K, order, w.Dr.data(), w.Ds.data(), w.L.data(), cx, cy, nbc, nbf, w.r.data(), w.s.data(),
                         1.0, 0.5, bc, v.data(), 0.2, out.data());
  for (int tau = 0; tau < 2; ++tau) {
    a = v;
    oracle_laplace_apply(K, order, st.Dr.data(), st.Ds.data(), st.L.data(), cx, cy, nbc, nbf, tau, a.data());
    for (int neu = 0; neu < 2; ++neu) {
      a = v;
      oracle_laplace_apply_bc(K, order, st.Dr.data(), st.Ds.data(), st.L.data(), cx, cy, nbc, nbf, tau, neu,
                              a.data());
      a = v;
      oracle_laplace_apply_data(K, order, st.Dr.data(), st.Ds.data(), st.L.data(), cx, cy, nbc, nbf, tau, neu,
                                g.data(), g.data(), a.data());
      a = v;
      oracle_laplace_inhomogeneous(K, order, st.Dr.data(), st.Ds.data(), st.L.data(), cx, cy, nbc, nbf, tau, neu,
                                   g.data(), g.data(), a.data());
    }
  }
  oracle_grad(K, order, w.Dr.data(), w.Ds.data(), w.L.data(), cx, cy, nbc, nbf, g.data(), -1.0, v.data(), a.data(),
              b.data());
  oracle_grad(K, order, w.Dr.data(), w.Ds.data(), w.L.data(), cx, cy, nbc, nbf, nullptr, 1.0, v.data(), a.data(),
              b.data());
  oracle_div(K, order, w.Dr.data(), w.Ds.data(), w.L.data(), cx, cy, nbc, nbf, g.data(), g.data(), -1.0, v.data(),
             v.data(), a.data());
  std::vector<double> st3(u);
  const double dt = 0.1 * c.min_edge / double((order + 1) * (order + 1));
  CHECK(oracle_rk4(0, K, order, w.Dr.data(), w.Ds.data(), w.L.data(), cx, cy, nbc, nbf, w.r.data(), w.s.data(), 1.0,
                   1.0, 0, st3.data(), 0.0, dt, 3, -1.0) == 3);
  a = v;
  CHECK(oracle_rk4(1, K, order, w.Dr.data(), w.Ds.data(), w.L.data(), cx, cy, nbc, nbf, w.r.data(), w.s.data(), 1.0,
                   0.5, 0, a.data(), 0.0, dt, 0, 2.5 * dt) == 3);
  for (double z : st3) CHECK(std::isfinite(z));
}

int main(int argc, char** argv) {
  const std::string golden = argc > 1 ? argv[1] : "tests/golden";
  const std::string tmp = argc > 2 ? argv[2] : "/tmp";
  for (int order = 1; order <= 7; ++order) {
    dgc_mesh_t m = nullptr;
    CHECK(dgc_mesh_structured(5, 4, -1, 1, -1, 1, 0.1, 7, 0, &m) == DGC_OK);
    exercise_mesh(m, order);
    dgc_mesh_destroy(m);
    CHECK(dgc_mesh_structured(6, 5, 0, 1, 0, 1, 0.05, 1, 1, &m) == DGC_OK);  // periodic
    exercise_mesh(m, order);
    dgc_mesh_destroy(m);
  }
  for (const char* name : {"square_domain_4cells.msh", "square_domain_coarse.msh", "square_domain_fine.msh"}) {
    dgc_mesh_t m = nullptr;
    CHECK(dgc_mesh_read_gmsh((golden + "/" + name).c_str(), &m) == DGC_OK);
    if (!m) continue;
    exercise_mesh(m, 2);
    const std::string out = tmp + "/sanitize_" + name;
    CHECK(dgc_mesh_write_gmsh(m, out.c_str()) == DGC_OK);
    dgc_mesh_t back = nullptr;
    CHECK(dgc_mesh_read_gmsh(out.c_str(), &back) == DGC_OK);
    int k1 = 0, k2 = 0, nv = 0;
    double me = 0;
    dgc_mesh_info(m, &k1, &nv, &me);
    dgc_mesh_info(back, &k2, &nv, &me);
    CHECK(k1 == k2);
    dgc_mesh_destroy(back);
    dgc_mesh_destroy(m);
    std::remove(out.c_str());
  }
  // mesh from arrays, plot files, the 1D oracle
  {
    const double vx[4] = {0, 1, 0, 1}, vy[4] = {0, 0, 1, 1};
    const int etov[6] = {0, 1, 2, 1, 3, 2};
    dgc_mesh_t m = nullptr;
    CHECK(dgc_mesh_from_arrays(4, vx, vy, 2, etov, &m) == DGC_OK);
    exercise_mesh(m, 3);
    dgc_mesh_destroy(m);
    const double xs[3] = {0, 0.5, 1}, us[3] = {1, 2, 3};
    const std::string p2 = tmp + "/sanitize_plot2d.txt", p1 = tmp + "/sanitize_plot1d.txt";
    CHECK(dgc_write_plot_file(p2.c_str(), 3, xs, xs, us, us) == DGC_OK);
    CHECK(dgc_write_plot_file(p1.c_str(), 3, xs, nullptr, us, us) == DGC_OK);
    std::remove(p2.c_str());
    std::remove(p1.c_str());
    for (int order = 1; order <= 8; ++order) {
      // the 1D oracle's memory use (operators: any (order+1)^2 matrices)
      const int K = 16, m1 = order + 1, n = K * m1;
      std::vector<double> D(m1 * m1), L(m1 * m1), u(n), o(n);
      for (int i = 0; i < m1 * m1; ++i) {
        D[i] = 0.01 * std::cos(double(i));
        L[i] = 0.01 * std::sin(double(i));
      }
      for (int i = 0; i < n; ++i) u[i] = std::sin(0.1 * i);
      oracle_adv1d_rhs(K, order, D.data(), L.data(), u.data(), 0.0, o.data());
      oracle_adv1d_rk4(K, order, D.data(), L.data(), u.data(), 0.0, 1e-3, 5);
    }
  }
  // error paths: bad arguments are reported, not dereferenced
  {
    dgc_mesh_t m = nullptr;
    CHECK(dgc_mesh_structured(0, 4, -1, 1, -1, 1, 0.1, 0, 0, &m) != DGC_OK);
    CHECK(dgc_mesh_structured(4, 4, 1, -1, -1, 1, 0.1, 0, 0, &m) != DGC_OK);
    CHECK(dgc_mesh_structured(4, 4, -1, 1, -1, 1, 0.4, 0, 0, &m) != DGC_OK);
    CHECK(dgc_mesh_read_gmsh((tmp + "/sanitize_no_such_file.msh").c_str(), &m) != DGC_OK);
    CHECK(dgc_ref_operators(0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) != DGC_OK);
    CHECK(dgc_ref_operators(3, 2, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) != DGC_OK);
    const double vx[3] = {0, 1, 0}, vy[3] = {0, 0, 1};
    const int bad[3] = {0, 1, 7};
    CHECK(dgc_mesh_from_arrays(3, vx, vy, 1, bad, &m) != DGC_OK);
    const int cw[3] = {0, 2, 1};  // clockwise cell
    dgc_mesh_t m2 = nullptr;
    const int rc = dgc_mesh_from_arrays(3, vx, vy, 1, cw, &m2);
    if (rc == DGC_OK) {
      int nb = 0;
      std::vector<double> inv(4), J(1), fJ(3), nx(3), ny(3);
      std::vector<int> ic(3), ifc(3);
      CHECK(dgc_mesh_geometry(m2, 2, inv.data(), J.data(), fJ.data(), nx.data(), ny.data(), ic.data(), ifc.data(),
                              &nb, nullptr, nullptr) != DGC_OK || !(J[0] > 0));
      dgc_mesh_destroy(m2);
    }
  }
  std::printf("host sanitizer run: %d check failures\n", failures);
  return failures == 0 ? 0 : 1;
}
