// Drop-in host C++ layer over the C ABI (include/dgc_hip.h).
//
// Mirrors the reference's host-facing names and call shapes
// (src/config.h:31-32 namespace macros). The CPU/GPU switch of the reference
// (USE_CPU_ONLY, src/config.h:29) is not a macro here: this layer always
// drives the gfx950 kernels, and the serial CPU path lives in the oracle.
#ifndef DGC_CONFIG_H
#define DGC_CONFIG_H

#include <stdexcept>
#include <string>

#include "dgc_hip.h"

#define BEGIN_NAMESPACE namespace dgc {
#define END_NAMESPACE }

BEGIN_NAMESPACE

// Status -> exception, with the reference's launch-failure texts where they
// exist (src/device_SemiDiscOp_wrapper.cuh:56, src/kernel_adapter.cuh:74,81).
inline void check(int rc, const char* what) {
  if (rc != DGC_OK) throw std::runtime_error(std::string(what) + " [" + dgc_last_error() + "]");
}

// Process-wide context on device 0, standing in for the legacy default stream.
inline dgc_ctx_t default_context() {
  static dgc_ctx_t ctx = [] {
    dgc_ctx_t c = nullptr;
    check(dgc_ctx_create(0, &c), "failed to create the device context");
    return c;
  }();
  return ctx;
}

END_NAMESPACE

#endif
