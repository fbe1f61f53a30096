// Poisson driver against the drop-in C++ layer (examples/poisson_2d/main.cpp:58-236
// flow, with the assembly + direct solve replaced by block-Jacobi PCG):
// -Lap u = 2 pi^2 sin(pi x) sin(pi y) on [-1,1]^2, u = 0 on the boundary.
//
//   poisson_2d [order] [mesh.msh | structured] [tau_mode]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "dgc/device_vector.h"
#include "dgc/laplace_2d.h"
#include "dgc/simple_triangular_mesh_2d.h"

using Mesh = dgc::simple_triangular_mesh_2d<double, int>;

int main(int argc, char** argv) {
  const int order = argc > 1 ? std::atoi(argv[1]) : 3;
  const std::string src = argc > 2 ? argv[2] : "structured";
  const int tau_mode = argc > 3 ? std::atoi(argv[3]) : 1;
  auto mesh = src == "structured" ? Mesh::structured(32, 32, -1, 1, -1, 1, 0.1, 3) : Mesh::import_gmsh(src);
  dgc::laplace_2d<double, Mesh> lap(*mesh, order, tau_mode);
  const int n = lap.total_num_nodes();
  std::vector<double> x(n), y(n), f(n), exact(n);
  dgc::check(dgc_mesh_dof_coords(mesh->handle(), order, 1, x.data(), y.data()), "dof coords failed");
  for (int i = 0; i < n; ++i) {
    exact[i] = std::sin(M_PI * x[i]) * std::sin(M_PI * y[i]);
    f[i] = 2.0 * M_PI * M_PI * exact[i];
  }
  dgc::device_vector<double> df(f), du(n);
  auto t0 = std::chrono::steady_clock::now();
  const int it = lap.solve(df.data(), du.data(), 1e-10);
  auto t1 = std::chrono::steady_clock::now();
  const auto u = du.to_host();
  double err = 0.0, mx = 0.0;
  for (int i = 0; i < n; ++i) {
    err += (u[i] - exact[i]) * (u[i] - exact[i]);
    mx = std::fmax(mx, std::fabs(u[i] - exact[i]));
  }
  // consistency of the in-place operator: Lap_h u ~ -f in the interior
  dgc::device_vector<double> dv(u);
  lap(dv.data());
  std::printf("K = %d, N = %d, PCG iterations = %d, error norm = %.6e, max error = %.3e\n", int(mesh->num_cells()),
              order, it, err / n, mx);
  std::printf("time used: %.3f ms\n", std::chrono::duration<double, std::milli>(t1 - t0).count());
  return std::isfinite(err) ? 0 : 1;  // a solve that did not converge has thrown already
}
