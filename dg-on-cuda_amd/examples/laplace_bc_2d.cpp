// laplace_2d with reference-style BC objects through the drop-in C++ layer
// (dgc/laplace_2d.h): the full operator()(u, bc) = apply_homogeneous -
// inhomogeneous part (src/laplace_2d.h:121-128) of u = sin(x) cos(y) on a
// jittered 10 x 8 mesh of [-1, 1]^2, with Dirichlet data on every boundary face
// and then Neumann data; prints the 2-norm of both results. Then the Dirichlet
// operator assembled as a CSR matrix (laplace_2d::assemble): prints the largest
// |A u - Lap_h u| relative to max |Lap_h u|.
//
//   laplace_bc_2d [order]   (default 3)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "dgc/device_vector.h"
#include "dgc/laplace_2d.h"
#include "dgc/simple_triangular_mesh_2d.h"

using Mesh = dgc::simple_triangular_mesh_2d<double, int>;

struct dirichlet_bc {
  bool is_dirichlet(double, double) const { return true; }
  double exterior_val(double x, double y, double a) const { return 2.0 * std::sin(x) * std::cos(y) - a; }
  double exterior_grad_n(double, double, double g) const { return g; }
};
struct neumann_bc {
  bool is_dirichlet(double, double) const { return false; }
  double exterior_val(double, double, double a) const { return a; }
  double exterior_grad_n(double x, double y, double g) const { return 2.0 * (std::cos(x) * std::sin(y) + x * y) - g; }
};

int main(int argc, char** argv) {
  const int N = argc > 1 ? std::atoi(argv[1]) : 3;
  auto mesh = Mesh::structured(10, 8, -1, 1, -1, 1, 0.15, 7);
  dgc::laplace_2d<double, Mesh> lap(*mesh, N);
  const std::size_t n = lap.total_num_nodes();
  std::vector<double> x(n), y(n), u(n);
  dgc::check(dgc_mesh_dof_coords(mesh->handle(), N, 1, x.data(), y.data()), "dof coordinates failed");
  for (std::size_t i = 0; i < n; ++i) u[i] = std::sin(x[i]) * std::cos(y[i]);
  auto norm = [](const std::vector<double>& v) {
    double s = 0.0;
    for (double a : v) s += a * a;
    return std::sqrt(s);
  };
  dgc::device_vector<double> d(u);
  lap(d.data(), dirichlet_bc{});
  const double nd = norm(d.to_host());
  dgc::device_vector<double> e(u);
  lap(e.data(), neumann_bc{});
  const double nn = norm(e.to_host());
  std::printf("order = %d, dirichlet norm = %.17e, neumann norm = %.17e\n", N, nd, nn);
  std::vector<long long> rp;
  std::vector<int> ci;
  std::vector<double> va;
  lap.assemble(dirichlet_bc{}, rp, ci, va);
  dgc::device_vector<double> f(u);
  lap.apply_homogeneous(f.data(), dirichlet_bc{});
  const std::vector<double> lu = f.to_host();
  double dmax = 0.0, lmax = 0.0;
  for (std::size_t r = 0; r < n; ++r) {
    double s = 0.0;
    for (long long q = rp[r]; q < rp[r + 1]; ++q) s += va[q] * u[ci[q]];
    dmax = std::fmax(dmax, std::fabs(s - lu[r]));
    lmax = std::fmax(lmax, std::fabs(lu[r]));
  }
  std::printf("assembled: nnz = %lld, max |A u - Lap u| / max |Lap u| = %.3e\n", rp[n], dmax / lmax);
  return 0;
}
