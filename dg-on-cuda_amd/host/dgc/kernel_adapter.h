// Phase launcher with the members and calls of src/kernel_adapter.cuh:59-83,
// for the library's IPDG device object (the role of d_poisson_2d,
// examples/poisson_2d/d_poisson_2d.cuh:55-172).
#ifndef DGC_KERNEL_ADAPTER_H
#define DGC_KERNEL_ADAPTER_H

#include <cstddef>
#include <stdexcept>

#include "config.h"

BEGIN_NAMESPACE

struct phase_one {};
struct phase_two {};

// Device object of the IPDG Laplacian (owns one dgc_ipdg_t).
class device_laplace_op {
 public:
  explicit device_laplace_op(dgc_ipdg_t h) : m_h(h) {}
  device_laplace_op(const device_laplace_op&) = delete;
  device_laplace_op& operator=(const device_laplace_op&) = delete;
  ~device_laplace_op() { dgc_ipdg_destroy(m_h); }
  dgc_ipdg_t handle() const { return m_h; }

 private:
  dgc_ipdg_t m_h;
};

template <typename DeviceFunctor>
struct kernel_adapter {
  DeviceFunctor* Dobj;
  int GridSize;
  int BlockSize;
  std::size_t NumExecs;

  kernel_adapter(DeviceFunctor* obj, int grid_size, int block_size, std::size_t num_execs)
      : Dobj(obj), GridSize(grid_size), BlockSize(block_size), NumExecs(num_execs) {}

  // the whole in-place apply in one call
  void execute(double* inout) {
    if (dgc_ipdg_apply(Dobj->handle(), inout) != DGC_OK)
      throw std::runtime_error("failed to launch kernel of device function!");
  }

  void execute_phase(phase_one, double* inout) { launch(1, inout); }
  void execute_phase(phase_two, double* inout) { launch(2, inout); }

 private:
  void launch(int phase, double* inout) {
    if (dgc_ipdg_phase(Dobj->handle(), phase, inout) != DGC_OK)
      throw std::runtime_error("failed to launch kernel of device function phase!");
  }
};

END_NAMESPACE

#endif
