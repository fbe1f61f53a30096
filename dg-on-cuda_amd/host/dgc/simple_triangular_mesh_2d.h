// Host mesh with the query API of src/mesh/simple_triangular_mesh_2d.h:42-149,
// backed by the library's C++ set-up (dgc_mesh_*). fill_vertices reads each
// (x, y) pair in order (the reference's unsequenced read at :65 is not reproduced).
#ifndef DGC_SIMPLE_TRIANGULAR_MESH_2D_H
#define DGC_SIMPLE_TRIANGULAR_MESH_2D_H

#include <cstdint>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "config.h"

BEGIN_NAMESPACE

template <typename CT, typename IT>
class simple_triangular_mesh_2d {
 public:
  using index_type = IT;

  simple_triangular_mesh_2d() = default;

  template <class InputIt>
  void fill_vertices(InputIt first, InputIt last) {
    m_vx.clear();
    m_vy.clear();
    while (first != last) {
      const CT x = *first++;
      const CT y = *first++;
      m_vx.push_back(x);
      m_vy.push_back(y);
    }
  }

  template <class InputIt>
  void fill_connectivity(InputIt first, InputIt last) {
    m_etov.assign(first, last);
  }

  void build_topology() {
    dgc_mesh_t h = nullptr;
    check(dgc_mesh_from_arrays(int(m_vx.size()), m_vx.data(), m_vy.data(), int(m_etov.size() / 3),
                               m_etov.data(), &h),
          "build_topology failed");
    adopt(h);
  }

  // Generated mesh (the structured, jittered meshes of the benchmark problems).
  static std::unique_ptr<simple_triangular_mesh_2d> structured(int nx, int ny, double x0, double x1,
                                                               double y0, double y1, double jitter,
                                                               uint64_t seed, bool periodic = false) {
    dgc_mesh_t h = nullptr;
    check(dgc_mesh_structured(nx, ny, x0, x1, y0, y1, jitter, seed, periodic ? 1 : 0, &h),
          "structured mesh failed");
    auto m = std::make_unique<simple_triangular_mesh_2d>();
    m->adopt(h);
    return m;
  }

  // gmsh MSH 4.1 (src/mesh/gmsh_importer_exporter.h:49-173).
  static std::unique_ptr<simple_triangular_mesh_2d> import_gmsh(const std::string& path) {
    dgc_mesh_t h = nullptr;
    check(dgc_mesh_read_gmsh(path.c_str(), &h), "gmsh import failed");
    auto m = std::make_unique<simple_triangular_mesh_2d>();
    m->adopt(h);
    return m;
  }

  // gmsh MSH 4.1 export (src/mesh/gmsh_importer_exporter.h:175-213).
  void export_gmsh(const std::string& path) const {
    check(dgc_mesh_write_gmsh(m_handle.get(), path.c_str()), "gmsh export failed");
  }

  IT num_cells() const { return IT(m_nb_cell.size() / 3); }
  IT num_vertices() const { return IT(m_nv); }
  double min_edge_length() const { return m_hmin; }

  // (isBoundary, nbCell, nbFace) as src/mesh/simple_triangular_mesh_2d.h:116-121
  std::tuple<bool, IT, int> get_face_neighbor(IT ci, int face) const {
    const std::size_t h = 3 * std::size_t(ci) + face;
    return std::make_tuple(m_nb_cell[h] == ci, IT(m_nb_cell[h]), m_nb_face[h]);
  }

  // cell vertex coordinates, counter-clockwise
  void get_cell(IT ci, CT x[3], CT y[3]) const {
    for (int v = 0; v < 3; ++v) {
      x[v] = m_cx[3 * std::size_t(ci) + v];
      y[v] = m_cy[3 * std::size_t(ci) + v];
    }
  }

  dgc_mesh_t handle() const { return m_handle.get(); }

 private:
  struct deleter {
    void operator()(dgc_mesh_s* m) const { dgc_mesh_destroy(m); }
  };

  void adopt(dgc_mesh_t h) {
    m_handle.reset(h);
    int K = 0, nv = 0;
    check(dgc_mesh_info(h, &K, &nv, &m_hmin), "mesh info failed");
    m_nv = nv;
    m_cx.resize(3 * std::size_t(K));
    m_cy.resize(3 * std::size_t(K));
    m_nb_cell.resize(3 * std::size_t(K));
    m_nb_face.resize(3 * std::size_t(K));
    std::vector<int> etov(3 * std::size_t(K));
    check(dgc_mesh_cells(h, m_cx.data(), m_cy.data(), etov.data(), m_nb_cell.data(), m_nb_face.data()),
          "mesh cells failed");
  }

  std::vector<double> m_vx, m_vy;
  std::vector<int> m_etov;
  std::unique_ptr<dgc_mesh_s, deleter> m_handle;
  int m_nv = 0;
  double m_hmin = 0.0;
  std::vector<double> m_cx, m_cy;
  std::vector<int> m_nb_cell, m_nb_face;
};

END_NAMESPACE

#endif
