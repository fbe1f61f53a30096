// Host discretization with the API of src/simple_discretization_2d.h:42-218:
// face node lists, inverse Jacobians and the fill_* marshalling used to build
// the device object; plus the device-object factory of
// src/d_simple_discretization_2d.cuh:89-152 (returning a handle).
#ifndef DGC_SIMPLE_DISCRETIZATION_2D_H
#define DGC_SIMPLE_DISCRETIZATION_2D_H

#include <memory>
#include <vector>

#include "config.h"

BEGIN_NAMESPACE

template <typename T, typename M>
class simple_discretization_2d {
 public:
  using index_type = typename M::index_type;

  simple_discretization_2d(const M& mesh, int order) : m_mesh(&mesh), m_order(order) {
    const int Np = (order + 1) * (order + 2) / 2, Nfp = order + 1;
    std::vector<int> fm(3 * Nfp);
    check(dgc_ref_operators(order, 0, nullptr, nullptr, fm.data(), nullptr, nullptr, nullptr, nullptr),
          "reference element failed");
    F0_Nodes.assign(fm.begin(), fm.begin() + Nfp);
    F1_Nodes.assign(fm.begin() + Nfp, fm.begin() + 2 * Nfp);
    F2_Nodes.assign(fm.begin() + 2 * Nfp, fm.end());
    const std::size_t K = mesh.num_cells();
    Inv_Jacobians.resize(4 * K);
    m_J.resize(K);
    m_faceJ.resize(3 * K);
    m_nx.resize(3 * K);
    m_ny.resize(3 * K);
    m_icell.resize(3 * K);
    m_iface.resize(3 * K);
    check(dgc_mesh_geometry(mesh.handle(), order, Inv_Jacobians.data(), m_J.data(), m_faceJ.data(), m_nx.data(),
                            m_ny.data(), m_icell.data(), m_iface.data(), &m_nbnodes, nullptr, nullptr),
          "geometry failed");
    m_bx.resize(m_nbnodes);
    m_by.resize(m_nbnodes);
    check(dgc_mesh_geometry(mesh.handle(), order, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                            &m_nbnodes, m_bx.data(), m_by.data()),
          "geometry failed");
    m_Np = Np;
  }

  index_type total_num_nodes() const { return index_type(m_mesh->num_cells()) * m_Np; }
  int order() const { return m_order; }
  const M& mesh() const { return *m_mesh; }

  template <typename It1, typename It2, typename It3>
  void fill_cell_mappings(It1 it1, It2 it2, It3 it3) const {
    for (std::size_t c = 0; c < m_J.size(); ++c) {
      for (int k = 0; k < 4; ++k) *it1++ = Inv_Jacobians[4 * c + k];
      *it2++ = m_J[c];
      for (int e = 0; e < 3; ++e) *it3++ = m_faceJ[3 * c + e];
    }
  }
  template <typename It1, typename It2>
  void fill_cell_interfaces(It1 it1, It2 it2) const {
    for (std::size_t h = 0; h < m_icell.size(); ++h) {
      *it1++ = m_icell[h];
      *it2++ = m_iface[h];
    }
  }
  template <typename It1, typename It2>
  index_type fill_boundary_nodes(It1 it1, It2 it2) const {
    for (int i = 0; i < m_nbnodes; ++i) {
      *it1++ = m_bx[i];
      *it2++ = m_by[i];
    }
    return index_type(m_nbnodes);
  }
  template <typename It1, typename It2>
  void fill_outward_normals(It1 it1, It2 it2) const {
    for (std::size_t h = 0; h < m_nx.size(); ++h) {
      *it1++ = m_nx[h];
      *it2++ = m_ny[h];
    }
  }

  std::vector<int> F0_Nodes, F1_Nodes, F2_Nodes;
  std::vector<T> Inv_Jacobians;

 protected:
  const M* m_mesh;
  int m_order;
  int m_Np = 0;
  std::vector<double> m_J, m_faceJ, m_nx, m_ny, m_bx, m_by;
  std::vector<int> m_icell, m_iface;
  int m_nbnodes = 0;
};

// Device copy of the discretization: same inputs as the reference factory,
// the library keeps the device arrays and hands back one handle.
using device_discretization = std::unique_ptr<dgc_disc_s, int (*)(dgc_disc_t)>;

template <typename F, typename I>
device_discretization create_simple_discretization_2d_on_device(
    I num_cells, int order, const int* h_face0_nodes, const int* h_face1_nodes, const int* h_face2_nodes,
    const F* h_inv_jacobians, const F* h_Js, const F* h_face_Js, const I* h_interface_cells,
    const I* h_interface_faces, I num_boundary_nodes, const F* h_boundary_node_Xs, const F* h_boundary_node_Ys,
    const F* h_outward_normal_Xs, const F* h_outward_normal_Ys, dgc_ctx_t ctx = default_context()) {
  dgc_disc_t d = nullptr;
  check(dgc_disc2d_create(ctx, int(num_cells), order, h_face0_nodes, h_face1_nodes, h_face2_nodes, h_inv_jacobians,
                          h_Js, h_face_Js, h_interface_cells, h_interface_faces, int(num_boundary_nodes),
                          h_boundary_node_Xs, h_boundary_node_Ys, h_outward_normal_Xs, h_outward_normal_Ys, &d),
        "failed to create the discretization on the device");
  return device_discretization(d, &dgc_disc2d_destroy);
}

// Convenience: device copy straight from a host discretization.
template <typename T, typename M>
device_discretization create_on_device(const simple_discretization_2d<T, M>& h, dgc_ctx_t ctx = default_context()) {
  std::vector<double> ij, J, fJ, bx, by, nx, ny;
  std::vector<int> ic, iff;
  h.fill_cell_mappings(std::back_inserter(ij), std::back_inserter(J), std::back_inserter(fJ));
  h.fill_cell_interfaces(std::back_inserter(ic), std::back_inserter(iff));
  const int nb = int(h.fill_boundary_nodes(std::back_inserter(bx), std::back_inserter(by)));
  h.fill_outward_normals(std::back_inserter(nx), std::back_inserter(ny));
  if (bx.empty()) {
    bx.push_back(0.0);
    by.push_back(0.0);
  }
  return create_simple_discretization_2d_on_device<double, int>(
      int(J.size()), h.order(), h.F0_Nodes.data(), h.F1_Nodes.data(), h.F2_Nodes.data(), ij.data(), J.data(),
      fJ.data(), ic.data(), iff.data(), nb, bx.data(), by.data(), nx.data(), ny.data(), ctx);
}

// Device set-up from the mesh alone (dgc_disc2d_create_from_cells): geometry,
// connectivity and boundary nodes built on the GPU from the cell corners and
// vertex ids; the same arrays, bit for bit, as create_on_device, without the
// host marshalling of src/simple_discretization_2d.h:84-218 (meshes of millions
// of cells).
template <typename M>
device_discretization create_on_device_from_mesh(const M& mesh, int order, dgc_ctx_t ctx = default_context()) {
  const std::size_t n = 3 * std::size_t(mesh.num_cells());
  std::vector<double> cx(n), cy(n);
  std::vector<int> etov(n);
  check(dgc_mesh_cells(mesh.handle(), cx.data(), cy.data(), etov.data(), nullptr, nullptr), "mesh cells failed");
  dgc_disc_t d = nullptr;
  check(dgc_disc2d_create_from_cells(ctx, int(mesh.num_cells()), int(mesh.num_vertices()), order, cx.data(),
                                     cy.data(), etov.data(), &d),
        "device set-up failed");
  return device_discretization(d, &dgc_disc2d_destroy);
}

END_NAMESPACE

#endif
