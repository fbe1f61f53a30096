// Device buffers and field bundles (replace thrust::device_vector and the
// thrust::zip_iterator of 3 device vectors used by examples/maxwell_2d/main.cpp:103-197).
#ifndef DGC_DEVICE_VECTOR_H
#define DGC_DEVICE_VECTOR_H

#include <cstddef>
#include <utility>
#include <vector>

#include "config.h"

BEGIN_NAMESPACE

template <typename T>
class device_vector {
 public:
  device_vector() = default;
  explicit device_vector(std::size_t n, dgc_ctx_t ctx = default_context()) : m_ctx(ctx), m_size(n) {
    check(dgc_malloc(m_ctx, n * sizeof(T), reinterpret_cast<void**>(&m_ptr)), "device allocation failed");
  }
  device_vector(const std::vector<T>& h, dgc_ctx_t ctx = default_context()) : device_vector(h.size(), ctx) {
    copy_from(h);
  }
  device_vector(const device_vector&) = delete;
  device_vector& operator=(const device_vector&) = delete;
  device_vector(device_vector&& o) noexcept { swap(o); }
  device_vector& operator=(device_vector&& o) noexcept {
    swap(o);
    return *this;
  }
  ~device_vector() {
    if (m_ptr) dgc_free(m_ctx, m_ptr);
  }
  void swap(device_vector& o) noexcept {
    std::swap(m_ctx, o.m_ctx);
    std::swap(m_ptr, o.m_ptr);
    std::swap(m_size, o.m_size);
  }
  T* data() { return m_ptr; }
  const T* data() const { return m_ptr; }
  std::size_t size() const { return m_size; }
  void copy_from(const std::vector<T>& h) {
    if (h.size() != m_size) throw std::invalid_argument("device_vector::copy_from: size mismatch");
    check(dgc_memcpy_h2d(m_ctx, m_ptr, h.data(), m_size * sizeof(T)), "host-to-device copy failed");
  }
  std::vector<T> to_host() const {
    std::vector<T> h(m_size);
    check(dgc_memcpy_d2h(m_ctx, h.data(), m_ptr, m_size * sizeof(T)), "device-to-host copy failed");
    return h;
  }

 private:
  dgc_ctx_t m_ctx = nullptr;
  T* m_ptr = nullptr;
  std::size_t m_size = 0;
};

// F device field pointers handled as one state, like a zip of F device vectors.
template <int F>
struct device_fields {
  double* p[F];
  double* const* ptrs() const { return p; }
  const double* const* cptrs() const { return p; }
};

inline device_fields<1> make_fields(device_vector<double>& a) { return {{a.data()}}; }
inline device_fields<3> make_fields(device_vector<double>& a, device_vector<double>& b,
                                    device_vector<double>& c) {
  return {{a.data(), b.data(), c.data()}};
}

END_NAMESPACE

#endif
