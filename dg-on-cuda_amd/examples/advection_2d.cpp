// 2D advection driver in the shape of examples/advection_2d/main.cpp:58-190:
// the example's problem (a = (1, 1), u0 = sin(pi (x + 1)) sin(pi (y + 1)),
// analytic inflow data) on a gmsh mesh, totalTSs RK4 steps of
// dt = (2/3) (l / N) (l / (2 + sqrt 2)) with l = 2 / sqrt(K), then the
// mean-square error against the exact solution and the example's plot file.
// create_device_object + rk4_on_device become dgc_advection_create and
// dgc_rk4_steps (the whole loop replayed as one graph per chunk of steps).
//
//   advection_2d [order [mesh.msh [totalTSs [plot file]]]]   (default 1 square_domain_coarse.msh 10000)
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "dgc/device_vector.h"
#include "dgc/simple_discretization_2d.h"
#include "dgc/simple_triangular_mesh_2d.h"

using Mesh = dgc::simple_triangular_mesh_2d<double, int>;

int main(int argc, char** argv) {
  const int order = argc > 1 ? std::atoi(argv[1]) : 1;
  const std::string meshFile = argc > 2 ? argv[2] : "square_domain_coarse.msh";
  const int totalTSs = argc > 3 ? std::atoi(argv[3]) : 10000;
  const char* plot = argc > 4 ? argv[4] : nullptr;
  auto mesh = Mesh::import_gmsh(meshFile);
  const int numCells = int(mesh->num_cells());
  auto disc = dgc::create_on_device_from_mesh(*mesh, order);
  const int Np = (order + 1) * (order + 2) / 2, NF = 3 * (order + 1);
  std::vector<double> Dr(Np * Np), Ds(Np * Np), L(Np * NF);
  dgc::check(dgc_ref_operators(order, 0, nullptr, nullptr, nullptr, Dr.data(), Ds.data(), L.data(), nullptr),
             "reference operators failed");
  dgc_op_t op = nullptr;
  dgc::check(dgc_advection_create(disc.get(), Dr.data(), Ds.data(), L.data(), 1.0, 1.0, 0, &op),
             "failed to create the device object");
  const int numDOFs = numCells * Np;
  std::vector<double> x(numDOFs), y(numDOFs), v0(numDOFs);
  dgc::check(dgc_mesh_dof_coords(mesh->handle(), order, 1, x.data(), y.data()), "dof coordinates failed");
  for (int i = 0; i < numDOFs; ++i) v0[i] = std::sin((x[i] + 1.0) * M_PI) * std::sin((y[i] + 1.0) * M_PI);
  dgc::device_vector<double> v(v0), v1(numDOFs), v2(numDOFs), v3(numDOFs), v4(numDOFs), v5(numDOFs);
  const double l = 2.0 / std::sqrt(double(numCells));
  const double dt = (2.0 / 3.0) * (l / double(order)) * (l / (2.0 + std::sqrt(2.0)));
  double* state[1] = {v.data()};
  double* wk[5] = {v1.data(), v2.data(), v3.data(), v4.data(), v5.data()};
  auto t0 = std::chrono::system_clock::now();
  dgc::check(dgc_rk4_steps(op, state, std::size_t(numDOFs), 0.0, dt, totalTSs, wk), "RK4 failed");
  dgc::check(dgc_ctx_sync(dgc::default_context()), "sync failed");
  auto t1 = std::chrono::system_clock::now();
  // the time the example reaches by adding dt totalTSs times
  double t = 0.0;
  for (int i = 0; i < totalTSs; ++i) t += dt;
  const std::vector<double> u = v.to_host();
  std::vector<double> ref(numDOFs);
  double err = 0.0;
  for (int i = 0; i < numDOFs; ++i) {
    ref[i] = std::sin((x[i] + 1.0 - t) * M_PI) * std::sin((y[i] + 1.0 - t) * M_PI);
    err += (ref[i] - u[i]) * (ref[i] - u[i]);
  }
  std::printf("K = %d, N = %d, %d steps of dt = %.6e\n", numCells, order, totalTSs, dt);
  std::printf("T = %g, error norm = %.6e\n", t, err / numDOFs);
  std::printf("time used: %lld ms\n",
              (long long)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count());
  if (plot) dgc::check(dgc_write_plot_file(plot, numDOFs, x.data(), y.data(), u.data(), ref.data()), "plot failed");
  dgc_op_destroy(op);
  return 0;
}
