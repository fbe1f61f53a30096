// Maxwell TM problem: the host class of examples/maxwell_2d/maxwell_2d.h:36-176
// (weak-form Dr/Ds/LIFT, DOF initialisation, the analytic cavity mode) and the
// kernel-side seam of examples/maxwell_2d/k_maxwell_2d.cuh:41-62
// (create_device_object, rk4_on_device, destroy_device_object).
// DOFs use the device (node-major) layout: field[j * K + c].
#ifndef DGC_MAXWELL_2D_H
#define DGC_MAXWELL_2D_H

#include <cmath>
#include <vector>

#include "device_SemiDiscOp_wrapper.h"
#include "explicit_runge_kutta.h"
#include "simple_discretization_2d.h"

BEGIN_NAMESPACE

template <typename T, typename M>
class maxwell_2d : public simple_discretization_2d<T, M> {
 public:
  maxwell_2d(const M& mesh, int order) : simple_discretization_2d<T, M>(mesh, order) {
    const int Np = (order + 1) * (order + 2) / 2;
    m_Dr.resize(Np * Np);
    m_Ds.resize(Np * Np);
    m_L.resize(Np * 3 * (order + 1));
    m_r.resize(Np);
    m_s.resize(Np);
    check(dgc_ref_operators(order, 0, m_r.data(), m_s.data(), nullptr, m_Dr.data(), m_Ds.data(), m_L.data(),
                            nullptr),
          "reference operators failed");
  }

  // node positions and the initial cavity mode: Ez = sin(pi x) sin(pi y), H = 0
  void initialize_dofs(std::vector<T>& x, std::vector<T>& y, std::vector<T>& Hx, std::vector<T>& Hy,
                       std::vector<T>& Ez) const {
    positions(x, y);
    Hx.assign(x.size(), T(0));
    Hy.assign(x.size(), T(0));
    Ez.resize(x.size());
    for (std::size_t i = 0; i < x.size(); ++i) Ez[i] = std::sin(x[i] * M_PI) * std::sin(y[i] * M_PI);
  }

  void exact_solution(T t, std::vector<T>& Hx, std::vector<T>& Hy, std::vector<T>& Ez) const {
    std::vector<T> x, y;
    positions(x, y);
    const T s2 = std::sqrt(T(2));
    Hx.resize(x.size());
    Hy.resize(x.size());
    Ez.resize(x.size());
    for (std::size_t i = 0; i < x.size(); ++i) {
      Hx[i] = -std::sin(x[i] * M_PI) * std::cos(y[i] * M_PI) * std::sin(s2 * M_PI * t) / s2;
      Hy[i] = std::cos(x[i] * M_PI) * std::sin(y[i] * M_PI) * std::sin(s2 * M_PI * t) / s2;
      Ez[i] = std::sin(x[i] * M_PI) * std::sin(y[i] * M_PI) * std::cos(s2 * M_PI * t);
    }
  }

  std::vector<T> m_Dr, m_Ds, m_L;  // weak-form operators, row-major

 private:
  void positions(std::vector<T>& x, std::vector<T>& y) const {
    x.resize(this->total_num_nodes());
    y.resize(this->total_num_nodes());
    check(dgc_mesh_dof_coords(this->mesh().handle(), this->order(), 1, x.data(), y.data()), "dof coords failed");
  }
  std::vector<T> m_r, m_s;
};

using d_maxwell_2d = device_semi_disc_op;

inline d_maxwell_2d* create_device_object(dgc_disc_t disc, const double* dr, const double* ds, const double* l) {
  dgc_op_t h = nullptr;
  check(dgc_maxwell_tm_create(disc, dr, ds, l, &h), "failed to create the Maxwell device object");
  return new d_maxwell_2d(h);
}

inline void rk4_on_device(int gridSize, int blockSize, device_fields<3> inout, std::size_t size, double t, double dt,
                          d_maxwell_2d* d_op, device_fields<3> wk0, device_fields<3> wk1, device_fields<3> wk2,
                          device_fields<3> wk3, device_fields<3> wk4) {
  device_SemiDiscOp_wrapper<d_maxwell_2d> w{d_op, gridSize, blockSize};
  rk4(inout, size, t, dt, w, 0, wk0, wk1, wk2, wk3, wk4);
}

inline void destroy_device_object(d_maxwell_2d* device_obj) { delete device_obj; }

END_NAMESPACE

#endif
