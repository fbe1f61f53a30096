// Launcher of a device semi-discrete operator: same members and call shape as
// src/device_SemiDiscOp_wrapper.cuh:45-58. The device object is a library
// operator (Maxwell TM or advection); the grid/block members are accepted for
// source compatibility, the library picks its own tiling (64 cells per
// 256-thread workgroup).
#ifndef DGC_DEVICE_SEMIDISCOP_WRAPPER_H
#define DGC_DEVICE_SEMIDISCOP_WRAPPER_H

#include <cstddef>
#include <stdexcept>

#include "config.h"
#include "device_vector.h"

BEGIN_NAMESPACE

// A device object: owns one dgc_op_t (the role of d_maxwell_2d / d_advection_2d).
class device_semi_disc_op {
 public:
  explicit device_semi_disc_op(dgc_op_t h) : m_h(h) {}
  device_semi_disc_op(const device_semi_disc_op&) = delete;
  device_semi_disc_op& operator=(const device_semi_disc_op&) = delete;
  ~device_semi_disc_op() { dgc_op_destroy(m_h); }
  dgc_op_t handle() const { return m_h; }
  int num_fields() const { return dgc_op_num_fields(m_h); }

 private:
  dgc_op_t m_h;
};

template <typename DeviceSemiDiscOp>
struct device_SemiDiscOp_wrapper {
  DeviceSemiDiscOp* m_Dop;
  int m_GridSize;
  int m_BlockSize;

  template <typename T, int F>
  void operator()(device_fields<F> in_cbegin, std::size_t size, T t, device_fields<F> out_begin) const {
    if (dgc_op_rhs(m_Dop->handle(), in_cbegin.cptrs(), size, double(t), out_begin.ptrs()) != DGC_OK)
      throw std::runtime_error("failed to launch kernel/device code!");
  }
};

END_NAMESPACE

#endif
