// Classical RK4 with the signature of src/explicit_runge_kutta.h:33-34.
//
// The generic template keeps the reference's contract (4 calls of op, 6 of
// axpy, caller-owned work vectors). When op is a device operator wrapper the
// overload below runs the fused step instead: 4 kernels, each evaluating the
// right-hand side and applying its stage update in one pass (dgc_rk4_step);
// the arithmetic and its operand order are the same.
#ifndef DGC_EXPLICIT_RUNGE_KUTTA_H
#define DGC_EXPLICIT_RUNGE_KUTTA_H

#include <cstddef>
#include <stdexcept>

#include "config.h"
#include "device_SemiDiscOp_wrapper.h"
#include "device_vector.h"

BEGIN_NAMESPACE

template <typename Itr, typename T, typename DiscreteOp, typename Axpy>
void rk4(Itr inout, std::size_t size, T t, T dt, const DiscreteOp& op, const Axpy& axpy, Itr wk0, Itr wk1,
         Itr wk2, Itr wk3, Itr wk4) {
  const T half = T(0.5) * dt;
  op(inout, size, t, wk1);                // k1
  axpy(half, wk1, size, inout, wk0);
  op(wk0, size, t + half, wk2);           // k2
  axpy(half, wk2, size, inout, wk0);
  op(wk0, size, t + half, wk3);           // k3
  axpy(dt, wk3, size, inout, wk0);
  op(wk0, size, t + dt, wk4);             // k4
  axpy(T(2), wk2, size, wk1, wk0);        // k1 + 2 k2
  axpy(T(2), wk3, size, wk4, wk1);        // k4 + 2 k3
  axpy(dt / T(6), wk0, size, inout, wk2);
  axpy(dt / T(6), wk1, size, wk2, inout);
}

template <int F, typename T, typename DOp, typename Axpy>
void rk4(device_fields<F> inout, std::size_t size, T t, T dt, const device_SemiDiscOp_wrapper<DOp>& op,
         const Axpy& /*axpy*/, device_fields<F> wk0, device_fields<F> wk1, device_fields<F> wk2,
         device_fields<F> wk3, device_fields<F> wk4) {
  double* wk[5 * F];
  const device_fields<F>* w[5] = {&wk0, &wk1, &wk2, &wk3, &wk4};
  for (int i = 0; i < 5; ++i)
    for (int f = 0; f < F; ++f) wk[i * F + f] = w[i]->p[f];
  if (dgc_rk4_step(op.m_Dop->handle(), inout.ptrs(), size, double(t), double(dt), wk) != DGC_OK)
    throw std::runtime_error("failed to launch kernel/device code!");
}

END_NAMESPACE

#endif
