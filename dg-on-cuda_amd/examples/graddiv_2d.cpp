// Gradient and divergence driver written against the drop-in C++ layer
// (dgc/grad_2d.h, dgc/div_2d.h) with reference-style BC objects: on a jittered
// mesh of [-1, 1]^2 the weak gradient of a polynomial of degree N with its own
// Dirichlet values is exact, and so is the divergence; prints both errors.
//
//   graddiv_2d [order]   (default 3)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <utility>
#include <vector>

#include "dgc/device_vector.h"
#include "dgc/div_2d.h"
#include "dgc/grad_2d.h"
#include "dgc/simple_triangular_mesh_2d.h"

using Mesh = dgc::simple_triangular_mesh_2d<double, int>;

static int N = 3;
static double p(double x, double y) { return std::pow(x, N) + 2 * x * std::pow(y, N - 1) - 3 * std::pow(y, N) + 0.5; }
static double q(double x, double y) { return std::pow(y, N) - x * std::pow(y, N - 1); }

// Dirichlet values through the central flux: exterior = 2 g - interior
struct p_bc {
  double exterior_val(double x, double y, double a) const { return 2 * p(x, y) - a; }
};
struct pq_bc {
  std::pair<double, double> exterior_val(double x, double y, std::pair<double, double> a) const {
    return {2 * p(x, y) - a.first, 2 * q(x, y) - a.second};
  }
};

int main(int argc, char** argv) {
  N = argc > 1 ? std::atoi(argv[1]) : 3;
  auto mesh = Mesh::structured(12, 10, -1, 1, -1, 1, 0.15, 5);
  dgc::grad_2d<double, Mesh> grad(*mesh, N);
  dgc::div_2d<double, Mesh> div(*mesh, N);
  const std::size_t n = grad.total_num_nodes();
  std::vector<double> x(n), y(n), hp(n), hq(n);
  dgc::check(dgc_mesh_dof_coords(mesh->handle(), N, 1, x.data(), y.data()), "dof coordinates failed");
  for (std::size_t i = 0; i < n; ++i) {
    hp[i] = p(x[i], y[i]);
    hq[i] = q(x[i], y[i]);
  }
  dgc::device_vector<double> u0(hp), u1(n);
  grad(u0.data(), u1.data(), p_bc{});
  const auto gx = u0.to_host(), gy = u1.to_host();
  double eg = 0.0;
  for (std::size_t i = 0; i < n; ++i) {
    const double px = N * std::pow(x[i], N - 1) + 2 * std::pow(y[i], N - 1);
    const double py = (N > 1 ? 2 * (N - 1) * x[i] * std::pow(y[i], N - 2) : 0.0) - 3 * N * std::pow(y[i], N - 1);
    eg = std::max(eg, std::max(std::abs(gx[i] - px), std::abs(gy[i] - py)));
  }
  dgc::device_vector<double> v0(hp), v1(hq);
  div(v0.data(), v1.data(), pq_bc{});
  const auto d = v0.to_host();
  double ed = 0.0;
  for (std::size_t i = 0; i < n; ++i) {
    const double px = N * std::pow(x[i], N - 1) + 2 * std::pow(y[i], N - 1);
    const double qy = N * std::pow(y[i], N - 1) - (N > 1 ? (N - 1) * x[i] * std::pow(y[i], N - 2) : 0.0);
    ed = std::max(ed, std::abs(d[i] - (px + qy)));
  }
  std::printf("order = %d, cells = %d, grad max error = %.3e, div max error = %.3e\n", N, int(mesh->num_cells()), eg,
              ed);
  return eg < 1e-9 && ed < 1e-9 ? 0 : 1;
}
