// Gradient and divergence with central fluxes, with the host API of
// src/grad_2d.h:36-59 and src/div_2d.h:36-59: constructed from (mesh, order),
// applied in place as operator()(in0, in1, bc) -- here on device vectors
// (node-major) and by the gfx950 kernels of dgc_grad2d_apply / dgc_div2d_apply.
//
// The reference's BC object (exterior_val(x, y, interior)) is called on the
// host at every boundary node with interior values 0 and 1, which gives the
// affine form g + beta * interior the kernels take; a BC that is not affine
// with one beta for every node throws.
#ifndef DGC_GRAD_2D_H
#define DGC_GRAD_2D_H

#include <cmath>
#include <memory>
#include <utility>
#include <vector>

#include "device_vector.h"
#include "simple_discretization_2d.h"

BEGIN_NAMESPACE

namespace detail {

struct graddiv_handle {
  explicit graddiv_handle(dgc_graddiv_t h) : h(h) {}
  ~graddiv_handle() { dgc_graddiv_destroy(h); }
  graddiv_handle(const graddiv_handle&) = delete;
  graddiv_handle& operator=(const graddiv_handle&) = delete;
  dgc_graddiv_t h;
};

// Device discretization + weak-form operators shared by grad_2d and div_2d.
template <typename T, typename M>
class graddiv_base : public simple_discretization_2d<T, M> {
 protected:
  graddiv_base(const M& mesh, int order, dgc_ctx_t ctx)
      : simple_discretization_2d<T, M>(mesh, order), m_ctx(ctx), m_disc(create_on_device(*this, ctx)) {
    const int Np = (order + 1) * (order + 2) / 2;
    std::vector<double> Dr(Np * Np), Ds(Np * Np), L(Np * 3 * (order + 1));
    check(dgc_ref_operators(order, 0, nullptr, nullptr, nullptr, Dr.data(), Ds.data(), L.data(), nullptr),
          "reference operators failed");
    dgc_graddiv_t h = nullptr;
    check(dgc_graddiv_create(m_disc.get(), Dr.data(), Ds.data(), L.data(), &h), "failed to create grad/div");
    m_op = std::make_unique<graddiv_handle>(h);
  }

  // g (on the device) and beta of an affine scalar BC, from two host calls per boundary node
  template <typename F>
  double affine_bc(F exterior, device_vector<T>& g) const {
    const int nb = this->m_nbnodes;
    std::vector<T> h(nb);
    double beta = 1.0;
    for (int b = 0; b < nb; ++b) {
      const auto [g0, g1] = exterior(this->m_bx[b], this->m_by[b]);
      h[b] = g0;
      const double bb = g1 - g0;
      if (b == 0) beta = bb;
      if (std::abs(bb - beta) > 1e-12 * std::max(1.0, std::abs(beta)))
        throw std::invalid_argument("grad_2d/div_2d: the boundary condition is not affine with one beta");
    }
    if (nb > 0) g.copy_from(h);
    return beta;
  }

  device_vector<T>& scratch() const {
    if (!m_tmp) m_tmp = std::make_unique<device_vector<T>>(this->total_num_nodes(), m_ctx);
    return *m_tmp;
  }
  device_vector<T>& bdata(int i) const {
    if (!m_g[i]) m_g[i] = std::make_unique<device_vector<T>>(std::max(1, this->m_nbnodes), m_ctx);
    return *m_g[i];
  }

  dgc_ctx_t m_ctx;
  device_discretization m_disc;
  std::unique_ptr<graddiv_handle> m_op;
  mutable std::unique_ptr<device_vector<T>> m_tmp, m_g[2];
};

}  // namespace detail

template <typename T, typename M>
class grad_2d : public detail::graddiv_base<T, M> {
 public:
  grad_2d(const M& mesh, int order, dgc_ctx_t ctx = default_context())
      : detail::graddiv_base<T, M>(mesh, order, ctx) {}

  // in place: the input is in0; the x and y components are stored in in0 and in1
  // (src/grad_2d.h:152-187). BC2D::exterior_val(x, y, interior) -> exterior.
  template <typename BC2D>
  void operator()(T* d_in0, T* d_in1, const BC2D& bc) const {
    auto& g = this->bdata(0);
    const double beta = this->affine_bc(
        [&](double x, double y) { return std::make_pair(bc.exterior_val(x, y, T(0)), bc.exterior_val(x, y, T(1))); },
        g);
    auto& tmp = this->scratch();
    const std::size_t n = this->total_num_nodes();
    check(dgc_memcpy_d2d(this->m_ctx, tmp.data(), d_in0, n * sizeof(T)), "copy failed");
    check(dgc_grad2d_apply(this->m_op->h, tmp.data(), this->m_nbnodes > 0 ? g.data() : nullptr, beta, d_in0,
                           d_in1),
          "grad_2d failed");
  }
};

template <typename T, typename M>
class div_2d : public detail::graddiv_base<T, M> {
 public:
  div_2d(const M& mesh, int order, dgc_ctx_t ctx = default_context()) : detail::graddiv_base<T, M>(mesh, order, ctx) {}

  // in place: the inputs are (in0, in1); the divergence is stored in in0
  // (src/div_2d.h:154-188). BC2D::exterior_val(x, y, pair) -> pair.
  template <typename BC2D>
  void operator()(T* d_in0, T* d_in1, const BC2D& bc) const {
    auto& g0 = this->bdata(0);
    auto& g1 = this->bdata(1);
    const double b0 = this->affine_bc(
        [&](double x, double y) {
          return std::make_pair(bc.exterior_val(x, y, std::make_pair(T(0), T(0))).first,
                                bc.exterior_val(x, y, std::make_pair(T(1), T(1))).first);
        },
        g0);
    const double b1 = this->affine_bc(
        [&](double x, double y) {
          return std::make_pair(bc.exterior_val(x, y, std::make_pair(T(0), T(0))).second,
                                bc.exterior_val(x, y, std::make_pair(T(1), T(1))).second);
        },
        g1);
    if (this->m_nbnodes > 0 && std::abs(b0 - b1) > 1e-12 * std::max(1.0, std::abs(b0)))
      throw std::invalid_argument("div_2d: the two components need the same beta");
    auto& tmp = this->scratch();
    const std::size_t n = this->total_num_nodes();
    check(dgc_memcpy_d2d(this->m_ctx, tmp.data(), d_in0, n * sizeof(T)), "copy failed");
    const bool nb = this->m_nbnodes > 0;
    check(dgc_div2d_apply(this->m_op->h, tmp.data(), d_in1, nb ? g0.data() : nullptr, nb ? g1.data() : nullptr, b0,
                          d_in0),
          "div_2d failed");
  }
};

END_NAMESPACE

#endif
