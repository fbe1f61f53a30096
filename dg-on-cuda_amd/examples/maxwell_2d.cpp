// Maxwell TM cavity driver written against the drop-in C++ layer, following
// the flow of examples/maxwell_2d/main.cpp:65-231: mesh -> problem ->
// fill_* -> device object -> time loop of rk4_on_device -> error norm.
//
//   maxwell_2d [order] [mesh.msh | structured] [host | device]
//   (default: order 4, the 16x16 jittered cavity mesh, dt = 0.5 h_min / (N+1)^2,
//   T = 1, geometry and connectivity set up on the host; "device" builds them
//   on the GPU from the mesh cells)
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "dgc/axpy.h"
#include "dgc/device_vector.h"
#include "dgc/maxwell_2d.h"
#include "dgc/simple_triangular_mesh_2d.h"

using Mesh = dgc::simple_triangular_mesh_2d<double, int>;

int main(int argc, char** argv) {
  const int order = argc > 1 ? std::atoi(argv[1]) : 4;
  const std::string src = argc > 2 ? argv[2] : "structured";
  auto mesh = src == "structured" ? Mesh::structured(16, 16, -1, 1, -1, 1, 0.1, 0) : Mesh::import_gmsh(src);

  dgc::maxwell_2d<double, Mesh> op(*mesh, order);
  const std::string setup = argc > 3 ? argv[3] : "host";
  auto disc = setup == "device" ? dgc::create_on_device_from_mesh(*mesh, order) : dgc::create_on_device(op);
  dgc::d_maxwell_2d* dOp = dgc::create_device_object(disc.get(), op.m_Dr.data(), op.m_Ds.data(), op.m_L.data());

  const int n = op.total_num_nodes();
  std::vector<double> x, y, Hx0, Hy0, Ez0;
  op.initialize_dofs(x, y, Hx0, Hy0, Ez0);
  dgc::device_vector<double> hx(Hx0), hy(Hy0), ez(Ez0);
  std::vector<dgc::device_vector<double>> w;
  for (int i = 0; i < 15; ++i) w.emplace_back(n);
  auto u = dgc::make_fields(hx, hy, ez);
  dgc::device_fields<3> wk[5];
  for (int i = 0; i < 5; ++i) wk[i] = dgc::make_fields(w[3 * i], w[3 * i + 1], w[3 * i + 2]);

  const double T = 1.0;
  double t = 0.0;
  double dt = 0.5 * mesh->min_edge_length() / ((order + 1.0) * (order + 1.0));
  int steps = 0;
  auto t0 = std::chrono::steady_clock::now();
  while (t < T) {
    if (t + dt > T) dt = T - t;
    dgc::rk4_on_device(0, 0, u, n, t, dt, dOp, wk[0], wk[1], wk[2], wk[3], wk[4]);
    t += dt;
    ++steps;
  }
  dgc::check(dgc_ctx_sync(dgc::default_context()), "sync failed");
  auto t1 = std::chrono::steady_clock::now();

  std::vector<double> eHx, eHy, eEz;
  op.exact_solution(t, eHx, eHy, eEz);
  const auto Hx = hx.to_host(), Hy = hy.to_host(), Ez = ez.to_host();
  double err = 0.0;
  for (int i = 0; i < n; ++i)
    err += (eHx[i] - Hx[i]) * (eHx[i] - Hx[i]) + (eHy[i] - Hy[i]) * (eHy[i] - Hy[i]) +
           (eEz[i] - Ez[i]) * (eEz[i] - Ez[i]);
  std::printf("K = %d, N = %d, steps = %d, T = %.17g, error norm = %.6e\n", int(mesh->num_cells()), order, steps,
              t, err / n);
  std::printf("time used: %.3f ms\n", std::chrono::duration<double, std::milli>(t1 - t0).count());
  dgc::destroy_device_object(dOp);
  return (err / n < 1e-2) ? 0 : 1;  // sanity bound; tests check tighter values
}
