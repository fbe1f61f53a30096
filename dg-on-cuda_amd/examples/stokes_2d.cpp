// Stokes driver in the shape of examples/stokes_2d/main.cpp:58-196: the
// example's analytic problem on a gmsh mesh, u^{n-1} = u^n = exact at t = 0,
// dt = 1e-3 to T = 0.1 (the last step shortened to land on T), then the
// mean-square error of both velocity components against the exact solution.
// The time loop's TODO (main.cpp:170-172) is dgc_stokes_advance; with a box
// count the pressure solves are two-level (dgc_stokes_coarse_space).
//
//   stokes_2d [order [mesh.msh [boxes]]]   (default 1 square_domain_fine.msh 0)
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

namespace {

// the exact velocity of examples/stokes_2d/stokes_2d.h:297-298
constexpr double kA = 2.883356;
constexpr double kLambda = 1.0 + kA * kA;

void exact(const std::vector<double>& x, const std::vector<double>& y, double t, std::vector<double>& ux,
           std::vector<double>& uy) {
  const double e = std::exp(-kLambda * t);
  for (std::size_t i = 0; i < x.size(); ++i) {
    ux[i] = std::sin(x[i]) * (kA * std::sin(kA * y[i]) - std::cos(kA) * std::sinh(y[i])) * e;
    uy[i] = std::cos(x[i]) * (std::cos(kA * y[i]) + std::cos(kA) * std::cosh(y[i])) * e;
  }
}

}  // namespace

int main(int argc, char** argv) {
  const int order = argc > 1 ? std::atoi(argv[1]) : 1;
  const std::string meshFile = argc > 2 ? argv[2] : "square_domain_fine.msh";
  const int boxes = argc > 3 ? std::atoi(argv[3]) : 0;
  auto mesh = Mesh::import_gmsh(meshFile);
  auto disc = dgc::create_on_device_from_mesh(*mesh, order);
  dgc_stokes_t op = nullptr;
  dgc::check(dgc_stokes_create(disc.get(), &op), "failed to create the Stokes object");
  if (boxes > 0) {
    const std::size_t n3 = 3 * std::size_t(mesh->num_cells());
    std::vector<double> cx(n3), cy(n3);
    std::vector<int> etov(n3);
    dgc::check(dgc_mesh_cells(mesh->handle(), cx.data(), cy.data(), etov.data(), nullptr, nullptr), "cells failed");
    dgc::check(dgc_stokes_coarse_space(op, cx.data(), cy.data(), boxes, boxes), "coarse space failed");
  }
  const int numNodes = int(mesh->num_cells()) * (order + 1) * (order + 2) / 2;
  std::vector<double> x(numNodes), y(numNodes), ux(numNodes), uy(numNodes);
  dgc::check(dgc_mesh_dof_coords(mesh->handle(), order, 1, x.data(), y.data()), "dof coordinates failed");
  exact(x, y, 0.0, ux, uy);
  // it0 (next), it1 (previous), it2 (current) of the example, on the device
  dgc::device_vector<double> ux0(ux), uy0(uy), ux1(ux), uy1(uy), ux2(ux), uy2(uy);
  const double T = 0.1;
  double t_prev = 0.0, t = 0.0, dt = 0.001;
  int iters[3] = {0, 0, 0}, steps = 0;
  auto t0 = std::chrono::system_clock::now();
  while (t < T) {
    if (t + dt > T) dt = T - t;
    dgc::check(dgc_stokes_advance(op, ux1.data(), uy1.data(), t_prev, ux2.data(), uy2.data(), t, dt, 1e-11,
                                  ux0.data(), uy0.data(), iters),
               "time step failed");
    // previous <- current <- next, by exchanging the buffers
    std::swap(ux1, ux2);
    std::swap(uy1, uy2);
    std::swap(ux2, ux0);
    std::swap(uy2, uy0);
    t_prev = t;
    t += dt;
    ++steps;
  }
  dgc::check(dgc_ctx_sync(dgc::default_context()), "sync failed");
  auto t1 = std::chrono::system_clock::now();
  std::vector<double> rx(numNodes), ry(numNodes);
  exact(x, y, t, rx, ry);
  const std::vector<double> hx = ux2.to_host(), hy = uy2.to_host();
  double err = 0.0;
  for (int i = 0; i < numNodes; ++i) err += (rx[i] - hx[i]) * (rx[i] - hx[i]) + (ry[i] - hy[i]) * (ry[i] - hy[i]);
  std::printf("K = %d, N = %d, %d steps, last step iterations: pressure %d, ux %d, uy %d\n", int(mesh->num_cells()),
              order, steps, iters[0], iters[1], iters[2]);
  std::printf("T = %g, error norm = %.6e\n", t, err / numNodes);
  std::printf("time used: %lld ms\n",
              (long long)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count());
  dgc_stokes_destroy(op);
  return 0;
}
