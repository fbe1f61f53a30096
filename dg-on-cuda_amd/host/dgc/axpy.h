// axpy on device fields: out = a * x + y (src/axpy.h:112-122 for zip iterators
// of device vectors; src/k_axpy.cuh:46-59 for single pointers).
#ifndef DGC_AXPY_H
#define DGC_AXPY_H

#include <cstddef>
#include <stdexcept>

#include "config.h"
#include "device_vector.h"

BEGIN_NAMESPACE

template <typename T>
void k_axpy_auto(T a, const T* x, std::size_t n, const T* y, T* out) {
  if (dgc_axpy(default_context(), double(a), x, n, y, out) != DGC_OK)
    throw std::runtime_error("failed to launch kernel/device code!");
}

template <typename T>
void k_axpy(int /*grid_size*/, int /*block_size*/, T a, const T* x, std::size_t n, const T* y, T* out) {
  k_axpy_auto(a, x, n, y, out);
}

template <typename T, int F>
void axpy_n(T a, device_fields<F> x, std::size_t x_size, device_fields<F> y, device_fields<F> out) {
  for (int f = 0; f < F; ++f) k_axpy_auto<double>(double(a), x.p[f], x_size, y.p[f], out.p[f]);
}

END_NAMESPACE

#endif
