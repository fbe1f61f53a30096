// IPDG Laplacian with the host API of src/laplace_2d.h:44-81 and the
// Poisson problem of examples/poisson_2d/poisson_2d.h (homogeneous Dirichlet
// data). operator() applies Lap_h in place on a device vector (node-major);
// solve() replaces the unit-vector assembly + sparse direct solve of
// examples/poisson_2d/main.cpp:132-223 by block-Jacobi PCG on -M Lap_h.
#ifndef DGC_LAPLACE_2D_H
#define DGC_LAPLACE_2D_H

#include <memory>
#include <vector>

#include "device_vector.h"
#include "kernel_adapter.h"
#include "simple_discretization_2d.h"

BEGIN_NAMESPACE

template <typename T, typename M>
class laplace_2d : public simple_discretization_2d<T, M> {
 public:
  // tau_mode 0: tau = 1 as the reference; 1: (N+1)(N+2)/h per face
  laplace_2d(const M& mesh, int order, int tau_mode = 0, dgc_ctx_t ctx = default_context())
      : simple_discretization_2d<T, M>(mesh, order), m_ctx(ctx), m_disc(create_on_device(*this, ctx)) {
    const int Np = (order + 1) * (order + 2) / 2;
    m_Dr.resize(Np * Np);
    m_Ds.resize(Np * Np);
    m_L.resize(Np * 3 * (order + 1));
    m_M.resize(Np * Np);
    check(dgc_ref_operators(order, 1, nullptr, nullptr, nullptr, m_Dr.data(), m_Ds.data(), m_L.data(), m_M.data()),
          "reference operators failed");
    dgc_ipdg_t h = nullptr;
    check(dgc_ipdg_create(m_disc.get(), m_Dr.data(), m_Ds.data(), m_L.data(), m_M.data(), tau_mode, 1.0, &h),
          "failed to create the IPDG device object");
    m_op = std::make_unique<device_laplace_op>(h);
  }

  // in place: u <- Lap_h u (homogeneous Dirichlet data on every boundary face)
  void operator()(T* d_u) const {
    kernel_adapter<device_laplace_op> ka(m_op.get(), 0, 0, this->total_num_nodes());
    ka.execute_phase(phase_one{}, d_u);
    ka.execute_phase(phase_two{}, d_u);
  }

  // The BC-object API of src/laplace_2d.h:52-62 on device vectors. The BC
  // object (is_dirichlet, exterior_val, exterior_grad_n at (x, y)) is called
  // on the host at every boundary node; one kind for all boundary faces.
  template <typename BC2D>
  void apply_homogeneous(T* d_u, const BC2D& bc) const {
    set_kind(bc);
    check(dgc_ipdg_apply(m_op->handle(), d_u), "apply failed");
  }
  // d_u += the inhomogeneous boundary part (src/laplace_2d.h:253-330)
  template <typename BC2D>
  void apply_inhomogeneous_as_rhs(T* d_u, const BC2D& bc) const {
    const bool dirichlet = set_kind(bc);
    const int nb = this->m_nbnodes;
    if (nb == 0) return;
    std::vector<T> h(nb);
    for (int b = 0; b < nb; ++b)
      h[b] = dirichlet ? bc.exterior_val(this->m_bx[b], this->m_by[b], T(0))
                       : bc.exterior_grad_n(this->m_bx[b], this->m_by[b], T(0));
    device_vector<T> g(h, m_ctx);
    check(dgc_ipdg_apply_inhomogeneous(m_op->handle(), dirichlet ? g.data() : nullptr,
                                       dirichlet ? nullptr : g.data(), d_u),
          "inhomogeneous apply failed");
  }
  // in place: u <- Lap_h u with the BC's data (src/laplace_2d.h:121-128):
  // the homogeneous part minus the inhomogeneous one
  template <typename BC2D>
  void operator()(T* d_u, const BC2D& bc) const {
    apply_homogeneous(d_u, bc);
    device_vector<T> z(this->total_num_nodes(), m_ctx);
    check(dgc_memset_zero(m_ctx, z.data(), z.size() * sizeof(T)), "memset failed");
    apply_inhomogeneous_as_rhs(z.data(), bc);
    check(dgc_axpy(m_ctx, -1.0, z.data(), z.size(), d_u, d_u), "axpy failed");
  }

  // The homogeneous operator of the BC's kind as a CSR matrix over the
  // node-major DOFs: the "implicit" laplace_2d of examples/stokes_2d/stokes_2d.h:74-76
  // in place of the unit-vector loop of stokes_2d.h:388-422.
  template <typename BC2D>
  void assemble(const BC2D& bc, std::vector<long long>& row_ptr, std::vector<int>& col_idx,
                std::vector<T>& values) const {
    set_kind(bc);
    long long nnz = 0;
    row_ptr.assign(this->total_num_nodes() + 1, 0);
    check(dgc_ipdg_assemble(m_op->handle(), &nnz, row_ptr.data(), nullptr, nullptr), "assembly sizes failed");
    col_idx.resize(std::size_t(nnz));
    values.resize(std::size_t(nnz));
    check(dgc_ipdg_assemble(m_op->handle(), &nnz, row_ptr.data(), col_idx.data(), values.data()),
          "assembly failed");
  }

  // -Lap_h u = f: b = M f, then block-Jacobi PCG to the relative residual rtol.
  // Returns the iteration count.
  int solve(const T* d_f, T* d_u, double rtol = 1e-10, int maxit = 100000) const {
    device_vector<T> b(this->total_num_nodes());
    check(dgc_ipdg_mass(m_op->handle(), d_f, b.data()), "mass apply failed");
    int it = 0;
    double rel = 0.0;
    float ms = 0.f;
    check(dgc_ipdg_pcg(m_op->handle(), b.data(), d_u, rtol, maxit, 1, &it, &rel, nullptr, &ms), "PCG failed");
    return it;
  }

  dgc_ipdg_t handle() const { return m_op->handle(); }

 private:
  // boundary kind of the BC object (one kind on every boundary node); true: Dirichlet
  template <typename BC2D>
  bool set_kind(const BC2D& bc) const {
    bool dirichlet = true;
    for (int b = 0; b < this->m_nbnodes; ++b) {
      const bool d = bc.is_dirichlet(this->m_bx[b], this->m_by[b]);
      if (b == 0) dirichlet = d;
      if (d != dirichlet) throw std::invalid_argument("laplace_2d: mixed Dirichlet/Neumann boundaries");
    }
    check(dgc_ipdg_set_boundary(m_op->handle(), dirichlet ? DGC_BC_DIRICHLET : DGC_BC_NEUMANN),
          "boundary kind failed");
    return dirichlet;
  }

  dgc_ctx_t m_ctx;
  device_discretization m_disc;
  std::unique_ptr<device_laplace_op> m_op;
  std::vector<T> m_Dr, m_Ds, m_L, m_M;
};

END_NAMESPACE

#endif
