// 1D advection driver in the shape of examples/advection_1d/main.cpp:55-147:
// operators, device object, RK4 loop, error norm, plot file. The host class
// advection_1d's m_D / m_L come from dgc_adv1d_operators; create_device_object
// and rk4_on_device become dgc_adv1d_create and dgc_adv1d_rk4_steps.
//
//   advection_1d [numCells order [totalTSs [plot file]]]   (default 8192 6 10000)
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "dgc/device_vector.h"

int main(int argc, char** argv) {
  const int numCells = argc > 2 ? std::atoi(argv[1]) : 1024 * 8;
  const int order = argc > 2 ? std::atoi(argv[2]) : 6;
  const int totalTSs = argc > 3 ? std::atoi(argv[3]) : 10000;
  const char* plot = argc > 4 ? argv[4] : nullptr;
  const int np = order + 1, numDOFs = numCells * np;
  std::vector<double> D(np * np), L(np * np), x(numDOFs);
  dgc::check(dgc_adv1d_operators(numCells, order, D.data(), L.data(), x.data()), "operators failed");
  dgc_adv1d_t op = nullptr;
  dgc::check(dgc_adv1d_create(dgc::default_context(), numCells, order, D.data(), L.data(), &op),
             "failed to create the device object");
  std::vector<double> v0(numDOFs);
  for (int i = 0; i < numDOFs; ++i) v0[i] = std::sin(x[i]);
  dgc::device_vector<double> v(v0);
  const double h = 2.0 * M_PI / numCells, a = 2.0 * M_PI;
  const double dt = 0.25 / order / order * h / a;
  auto t0 = std::chrono::system_clock::now();
  dgc::check(dgc_adv1d_rk4_steps(op, v.data(), 0.0, dt, totalTSs), "RK4 failed");
  dgc::check(dgc_ctx_sync(dgc::default_context()), "sync failed");
  auto t1 = std::chrono::system_clock::now();
  const double t = totalTSs * dt;
  const std::vector<double> u = v.to_host();
  std::vector<double> ref(numDOFs);
  double err = 0.0;
  for (int i = 0; i < numDOFs; ++i) {
    ref[i] = std::sin(x[i] - a * t);
    err += (ref[i] - u[i]) * (ref[i] - u[i]);
  }
  std::printf("t = %g, error norm = %g\n", t, err / numDOFs);
  std::printf("time used: %lld ms\n",
              (long long)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count());
  if (plot) dgc::check(dgc_write_plot_file(plot, numDOFs, x.data(), nullptr, u.data(), ref.data()), "plot failed");
  dgc_adv1d_destroy(op);
  return 0;
}
