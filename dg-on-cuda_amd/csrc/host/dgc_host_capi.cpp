// C ABI of the host set-up: reference-element operators, meshes, geometry.
#include <cstring>
#include <exception>
#include <fstream>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/dgc_hip.h"
#include "mesh.hpp"
#include "refelem.hpp"

int dgc_set_error(int code, const std::string& msg);

struct dgc_mesh_s {
  dgc_host::Mesh2D mesh;
};

namespace {
// Copies a host vector to a caller array (nothing for an empty vector, whose
// data() may be null: memcpy from a null pointer is undefined even for 0 bytes).
template <class T>
void copy_out(T* dst, const std::vector<T>& v, size_t count) {
  if (dst != nullptr && count > 0) std::memcpy(dst, v.data(), count * sizeof(T));
}

template <class Fn>
int guarded(const char* what, Fn&& fn) {
  try {
    return fn();
  } catch (const std::exception& e) {
    return dgc_set_error(DGC_ERR_INVALID, std::string(what) + ": " + e.what());
  }
}
}  // namespace

extern "C" {

int dgc_ref_operators(int order, int form, double* r, double* s, int* fmask, double* Dr, double* Ds,
                      double* lift, double* mass) {
  return guarded("dgc_ref_operators", [&]() -> int {
    if (order < 1 || order > 15) return dgc_set_error(DGC_ERR_INVALID, "dgc_ref_operators: order out of range");
    if (form != 0 && form != 1) return dgc_set_error(DGC_ERR_INVALID, "dgc_ref_operators: form must be 0 or 1");
    dgc_host::RefTriangle ref(order, form == 0 ? dgc_host::kWeakForm : dgc_host::kStrongForm);
    if (r) std::memcpy(r, ref.r.data(), ref.r.size() * sizeof(double));
    if (s) std::memcpy(s, ref.s.data(), ref.s.size() * sizeof(double));
    if (fmask) std::memcpy(fmask, ref.fmask.data(), ref.fmask.size() * sizeof(int));
    if (Dr) std::memcpy(Dr, ref.Dr.a.data(), ref.Dr.a.size() * sizeof(double));
    if (Ds) std::memcpy(Ds, ref.Ds.a.data(), ref.Ds.a.size() * sizeof(double));
    if (lift) std::memcpy(lift, ref.LIFT.a.data(), ref.LIFT.a.size() * sizeof(double));
    if (mass) std::memcpy(mass, ref.Mass.a.data(), ref.Mass.a.size() * sizeof(double));
    return DGC_OK;
  });
}

int dgc_mesh_structured(int nx, int ny, double x0, double x1, double y0, double y1, double jitter,
                        uint64_t seed, int periodic, dgc_mesh_t* out) {
  return guarded("dgc_mesh_structured", [&]() -> int {
    if (!out) return dgc_set_error(DGC_ERR_INVALID, "dgc_mesh_structured: out is null");
    if (!(x1 > x0) || !(y1 > y0)) return dgc_set_error(DGC_ERR_INVALID, "dgc_mesh_structured: empty domain");
    if (!(jitter >= 0.0 && jitter < 0.25))
      return dgc_set_error(DGC_ERR_INVALID, "dgc_mesh_structured: jitter must be in [0, 0.25)");
    if (double(nx) * ny * 2.0 >= double(1 << 29))
      return dgc_set_error(DGC_ERR_INVALID, "dgc_mesh_structured: too many cells");
    auto m = std::make_unique<dgc_mesh_s>();
    m->mesh = dgc_host::structured_mesh(nx, ny, x0, x1, y0, y1, jitter, seed, periodic != 0);
    *out = m.release();
    return DGC_OK;
  });
}

int dgc_mesh_from_arrays(int nv, const double* vx, const double* vy, int k, const int* etov,
                         dgc_mesh_t* out) {
  return guarded("dgc_mesh_from_arrays", [&]() -> int {
    if (!out || !vx || !vy || !etov || nv < 3 || k < 1)
      return dgc_set_error(DGC_ERR_INVALID, "dgc_mesh_from_arrays: bad argument");
    auto m = std::make_unique<dgc_mesh_s>();
    m->mesh = dgc_host::mesh_from_arrays(nv, vx, vy, k, etov);
    *out = m.release();
    return DGC_OK;
  });
}

int dgc_mesh_read_gmsh(const char* path, dgc_mesh_t* out) {
  return guarded("dgc_mesh_read_gmsh", [&]() -> int {
    if (!out || !path) return dgc_set_error(DGC_ERR_INVALID, "dgc_mesh_read_gmsh: null argument");
    std::vector<double> vx, vy;
    std::vector<int> etov;
    std::string err;
    if (!dgc_host::read_gmsh41(path, vx, vy, etov, err))
      return dgc_set_error(DGC_ERR_INVALID, "dgc_mesh_read_gmsh: " + err);
    auto m = std::make_unique<dgc_mesh_s>();
    m->mesh = dgc_host::mesh_from_arrays(int(vx.size()), vx.data(), vy.data(), int(etov.size() / 3),
                                         etov.data());
    *out = m.release();
    return DGC_OK;
  });
}

int dgc_mesh_write_gmsh(dgc_mesh_t mesh, const char* path) {
  return guarded("dgc_mesh_write_gmsh", [&]() -> int {
    if (!mesh || !path) return dgc_set_error(DGC_ERR_INVALID, "dgc_mesh_write_gmsh: null argument");
    std::string err;
    if (!dgc_host::write_gmsh41(mesh->mesh, path, err))
      return dgc_set_error(DGC_ERR_INVALID, "dgc_mesh_write_gmsh: " + err);
    return DGC_OK;
  });
}

int dgc_mesh_destroy(dgc_mesh_t mesh) {
  delete mesh;
  return DGC_OK;
}

int dgc_mesh_info(dgc_mesh_t mesh, int* num_cells, int* num_vertices, double* min_edge) {
  if (!mesh) return dgc_set_error(DGC_ERR_INVALID, "dgc_mesh_info: null mesh");
  if (num_cells) *num_cells = mesh->mesh.K;
  if (num_vertices) *num_vertices = mesh->mesh.Nv;
  if (min_edge) *min_edge = dgc_host::min_edge_length(mesh->mesh);
  return DGC_OK;
}

int dgc_mesh_cells(dgc_mesh_t mesh, double* cx, double* cy, int* etov, int* nb_cell, int* nb_face) {
  if (!mesh) return dgc_set_error(DGC_ERR_INVALID, "dgc_mesh_cells: null mesh");
  const auto& m = mesh->mesh;
  const size_t n = size_t(3) * m.K;
  if (cx) std::memcpy(cx, m.cx.data(), n * sizeof(double));
  if (cy) std::memcpy(cy, m.cy.data(), n * sizeof(double));
  if (etov) std::memcpy(etov, m.etov.data(), n * sizeof(int));
  if (nb_cell) std::memcpy(nb_cell, m.nb_cell.data(), n * sizeof(int));
  if (nb_face) std::memcpy(nb_face, m.nb_face.data(), n * sizeof(int));
  return DGC_OK;
}

int dgc_mesh_geometry(dgc_mesh_t mesh, int order, double* inv_jac, double* J, double* face_J,
                      double* nx, double* ny, int* iface_cell, int* iface_face,
                      int* num_boundary_nodes, double* bx, double* by) {
  return guarded("dgc_mesh_geometry", [&]() -> int {
    if (!mesh) return dgc_set_error(DGC_ERR_INVALID, "dgc_mesh_geometry: null mesh");
    if (order < 1 || order > 15) return dgc_set_error(DGC_ERR_INVALID, "dgc_mesh_geometry: order out of range");
    std::vector<double> r, s;
    dgc_host::triangle_nodes(order, r, s);
    const dgc_host::Geometry g = dgc_host::build_geometry(mesh->mesh, order, r, s);
    const size_t K = size_t(g.K);
    copy_out(inv_jac, g.inv_jac, 4 * K);
    copy_out(J, g.J, K);
    copy_out(face_J, g.face_J, 3 * K);
    copy_out(nx, g.nx, 3 * K);
    copy_out(ny, g.ny, 3 * K);
    copy_out(iface_cell, g.iface_cell, 3 * K);
    copy_out(iface_face, g.iface_face, 3 * K);
    if (num_boundary_nodes) *num_boundary_nodes = g.num_boundary_nodes;
    copy_out(bx, g.bnode_x, g.bnode_x.size());
    copy_out(by, g.bnode_y, g.bnode_y.size());
    return DGC_OK;
  });
}

int dgc_mesh_dof_coords(dgc_mesh_t mesh, int order, int node_major, double* x, double* y) {
  return guarded("dgc_mesh_dof_coords", [&]() -> int {
    if (!mesh || !x || !y) return dgc_set_error(DGC_ERR_INVALID, "dgc_mesh_dof_coords: null argument");
    std::vector<double> r, s, xs, ys;
    dgc_host::triangle_nodes(order, r, s);
    dgc_host::dof_coordinates(mesh->mesh, r, s, node_major != 0, xs, ys);
    std::memcpy(x, xs.data(), xs.size() * sizeof(double));
    std::memcpy(y, ys.data(), ys.size() * sizeof(double));
    return DGC_OK;
  });
}

int dgc_write_plot_file(const char* path, size_t n, const double* x, const double* y, const double* u,
                        const double* ref) {
  if (!path || !x || !u || !ref) return dgc_set_error(DGC_ERR_INVALID, "dgc_write_plot_file: null argument");
  std::ofstream file(path);
  if (!file) return dgc_set_error(DGC_ERR_INVALID, std::string("dgc_write_plot_file: cannot open ") + path);
  file.precision(std::numeric_limits<double>::digits10);
  if (y) {
    // examples/advection_2d/main.cpp:164-174
    file << "#         x         y         u" << std::endl;
    for (size_t i = 0; i < n; ++i) file << x[i] << "  " << y[i] << "  " << u[i] << std::endl;
    file << std::endl;
    file << "#         x         y         reference solution" << std::endl;
    for (size_t i = 0; i < n; ++i) file << x[i] << "  " << y[i] << "  " << ref[i] << std::endl;
  } else {
    // examples/advection_1d/main.cpp:124-134 (its reference block has one space)
    file << "#         x         y" << std::endl;
    for (size_t i = 0; i < n; ++i) file << x[i] << "  " << u[i] << std::endl;
    file << std::endl;
    file << "#         x         reference solution" << std::endl;
    for (size_t i = 0; i < n; ++i) file << x[i] << " " << ref[i] << std::endl;
  }
  if (!file) return dgc_set_error(DGC_ERR_INVALID, std::string("dgc_write_plot_file: write failed: ") + path);
  return DGC_OK;
}

}  // extern "C"
