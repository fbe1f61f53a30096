#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace dgc_host {

Mesh2D structured_mesh(int nx, int ny, double x0, double x1, double y0, double y1, double jitter,
                       uint64_t seed, bool periodic) {
  if (nx < 1 || ny < 1) throw std::invalid_argument("structured_mesh: nx, ny must be >= 1");
  if (periodic && (nx < 3 || ny < 3))
    throw std::invalid_argument("structured_mesh: periodic meshes need nx, ny >= 3");
  Mesh2D m;
  m.periodic = periodic;
  const double hx = (x1 - x0) / nx, hy = (y1 - y0) / ny;
  const double h = std::min(hx, hy);
  // Vertex grid (nx+1) x (ny+1); periodic meshes use only the first nx x ny.
  const int vnx = nx + 1, vny = ny + 1;
  std::vector<double> gx(size_t(vnx) * vny), gy(size_t(vnx) * vny);
  SplitMix64 rng(seed);
  for (int j = 0; j < vny; ++j)
    for (int i = 0; i < vnx; ++i) {
      const double dx = jitter * h * (2.0 * rng.uniform() - 1.0);
      const double dy = jitter * h * (2.0 * rng.uniform() - 1.0);
      const bool interior = (i > 0 && i < nx && j > 0 && j < ny);
      double x = (i == nx) ? x1 : x0 + i * hx;
      double y = (j == ny) ? y1 : y0 + j * hy;
      if (periodic || interior) {
        x += dx;
        y += dy;
      }
      gx[size_t(j) * vnx + i] = x;
      gy[size_t(j) * vnx + i] = y;
    }
  if (periodic) {
    // the wrap vertices are images of the first row / column
    for (int j = 0; j < vny; ++j)
      for (int i = 0; i < vnx; ++i) {
        if (i < nx && j < ny) continue;
        const int si = i % nx, sj = j % ny;
        gx[size_t(j) * vnx + i] = gx[size_t(sj) * vnx + si] + (i == nx ? (x1 - x0) : 0.0);
        gy[size_t(j) * vnx + i] = gy[size_t(sj) * vnx + si] + (j == ny ? (y1 - y0) : 0.0);
      }
    m.Nv = nx * ny;
    m.px.resize(m.Nv);
    m.py.resize(m.Nv);
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        m.px[size_t(j) * nx + i] = gx[size_t(j) * vnx + i];
        m.py[size_t(j) * nx + i] = gy[size_t(j) * vnx + i];
      }
  } else {
    m.Nv = vnx * vny;
    m.px = gx;
    m.py = gy;
  }
  auto vid = [&](int i, int j) {
    return periodic ? (j % ny) * nx + (i % nx) : j * vnx + i;
  };
  m.K = 2 * nx * ny;
  m.etov.resize(size_t(3) * m.K);
  m.cx.resize(size_t(3) * m.K);
  m.cy.resize(size_t(3) * m.K);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int q = i + j * nx;
      const int ci[2][3] = {{i, i + 1, i + 1}, {i, i + 1, i}};
      const int cj[2][3] = {{j, j, j + 1}, {j, j + 1, j + 1}};
      for (int t = 0; t < 2; ++t) {
        const size_t c = size_t(2) * q + t;
        for (int v = 0; v < 3; ++v) {
          const int ii = ci[t][v], jj = cj[t][v];
          m.etov[3 * c + v] = vid(ii, jj);
          m.cx[3 * c + v] = gx[size_t(jj) * vnx + ii];
          m.cy[3 * c + v] = gy[size_t(jj) * vnx + ii];
        }
      }
    }
  build_topology(m);
  return m;
}

Mesh2D mesh_from_arrays(int nv, const double* vx, const double* vy, int k, const int* etov) {
  Mesh2D m;
  m.K = k;
  m.Nv = nv;
  m.px.assign(vx, vx + nv);
  m.py.assign(vy, vy + nv);
  m.etov.assign(etov, etov + size_t(3) * k);
  m.cx.resize(size_t(3) * k);
  m.cy.resize(size_t(3) * k);
  for (size_t i = 0; i < size_t(3) * k; ++i) {
    const int v = m.etov[i];
    if (v < 0 || v >= nv) throw std::out_of_range("mesh_from_arrays: vertex id out of range");
    m.cx[i] = vx[v];
    m.cy[i] = vy[v];
  }
  build_topology(m);
  return m;
}

void build_topology(Mesh2D& m) {
  // Sort half-edges by their (min, max) vertex key; equal neighbours are twins.
  const size_t nh = size_t(3) * m.K;
  std::vector<uint64_t> key(nh);
  std::vector<int> order(nh);
  for (size_t h = 0; h < nh; ++h) {
    const size_t c = h / 3;
    const int e = int(h % 3);
    uint64_t a = uint32_t(m.etov[3 * c + e]), b = uint32_t(m.etov[3 * c + (e + 1) % 3]);
    if (a == b) throw std::runtime_error("build_topology: degenerate edge");
    if (a > b) std::swap(a, b);
    key[h] = (a << 32) | b;
    order[h] = int(h);
  }
  std::sort(order.begin(), order.end(), [&](int p, int q) {
    return key[p] < key[q] || (key[p] == key[q] && p < q);
  });
  m.nb_cell.resize(nh);
  m.nb_face.resize(nh);
  for (size_t h = 0; h < nh; ++h) {
    m.nb_cell[h] = int(h / 3);
    m.nb_face[h] = int(h % 3);
  }
  for (size_t i = 0; i < nh;) {
    size_t j = i + 1;
    while (j < nh && key[order[j]] == key[order[i]]) ++j;
    if (j - i == 2) {
      const int p = order[i], q = order[i + 1];
      m.nb_cell[p] = q / 3;
      m.nb_face[p] = q % 3;
      m.nb_cell[q] = p / 3;
      m.nb_face[q] = p % 3;
    } else if (j - i > 2) {
      throw std::runtime_error("build_topology: non-manifold edge");
    }
    i = j;
  }
}

double min_edge_length(const Mesh2D& m) {
  double hmin = 1e300;
  for (int c = 0; c < m.K; ++c)
    for (int e = 0; e < 3; ++e) {
      const double dx = m.cx[3 * c + e] - m.cx[3 * c + (e + 1) % 3];
      const double dy = m.cy[3 * c + e] - m.cy[3 * c + (e + 1) % 3];
      hmin = std::min(hmin, std::sqrt(dx * dx + dy * dy));
    }
  return hmin;
}

Geometry build_geometry(const Mesh2D& m, int N, const std::vector<double>& r,
                        const std::vector<double>& s) {
  Geometry g;
  g.K = m.K;
  g.N = N;
  g.Nfp = N + 1;
  const int K = m.K;
  g.inv_jac.resize(size_t(4) * K);
  g.J.resize(K);
  g.face_J.resize(size_t(3) * K);
  g.nx.resize(size_t(3) * K);
  g.ny.resize(size_t(3) * K);
  g.iface_cell.resize(size_t(3) * K);
  g.iface_face.resize(size_t(3) * K);
  int offset = 0;
  for (int c = 0; c < K; ++c) {
    const double x1 = m.cx[3 * c], y1 = m.cy[3 * c];
    const double x2 = m.cx[3 * c + 1], y2 = m.cy[3 * c + 1];
    const double x3 = m.cx[3 * c + 2], y3 = m.cy[3 * c + 2];
    // [[x_r, x_s], [y_r, y_s]]
    const double xr = 0.5 * (x2 - x1), xs = 0.5 * (x3 - x1);
    const double yr = 0.5 * (y2 - y1), ys = 0.5 * (y3 - y1);
    const double det = xr * ys - xs * yr;
    if (det == 0.0) throw std::runtime_error("build_geometry: degenerate cell");
    g.inv_jac[4 * c + 0] = ys / det;   // rx
    g.inv_jac[4 * c + 1] = -xs / det;  // ry
    g.inv_jac[4 * c + 2] = -yr / det;  // sx
    g.inv_jac[4 * c + 3] = xr / det;   // sy
    g.J[c] = 0.25 * ((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1));
    const double vx[3] = {x1, x2, x3}, vy[3] = {y1, y2, y3};
    for (int e = 0; e < 3; ++e) {
      const int a = e, b = (e + 1) % 3, o = (e + 2) % 3;
      double ex = vx[b] - vx[a], ey = vy[b] - vy[a];
      g.face_J[3 * c + e] = 0.5 * std::sqrt(ex * ex + ey * ey);
      // (ey, -ex) is outward for a counter-clockwise cell; flip if it points at the third vertex
      const double ox = vx[o] - vx[a], oy = vy[o] - vy[a];
      if (ey * ox - ex * oy > 0.0) {
        ex = -ex;
        ey = -ey;
      }
      const double l = std::sqrt(ex * ex + ey * ey);
      g.nx[3 * c + e] = ey / l;
      g.ny[3 * c + e] = -ex / l;
      const int h = 3 * c + e;
      if (m.nb_cell[h] == c && m.nb_face[h] == e) {
        g.iface_cell[h] = c;
        g.iface_face[h] = offset;
        for (int d = 0; d < g.Nfp; ++d) {
          // rs -> xy with barycentric weights of vertices 1, 2, 3
          const int n = (e == 0) ? d : (e == 1 ? N + d * N - d * (d - 1) / 2
                                                : (N + 1) * (N + 2) / 2 - 1 - d * (d + 3) / 2);
          const double l1 = 0.5 * (s[n] + 1.0), l2 = -0.5 * (r[n] + s[n]), l3 = 0.5 * (r[n] + 1.0);
          g.bnode_x.push_back(l2 * x1 + l3 * x2 + l1 * x3);
          g.bnode_y.push_back(l2 * y1 + l3 * y2 + l1 * y3);
        }
        offset += g.Nfp;
      } else {
        g.iface_cell[h] = m.nb_cell[h];
        g.iface_face[h] = m.nb_face[h];
      }
    }
  }
  g.num_boundary_nodes = offset;
  return g;
}

void dof_coordinates(const Mesh2D& m, const std::vector<double>& r, const std::vector<double>& s,
                     bool node_major, std::vector<double>& x, std::vector<double>& y) {
  const int Np = int(r.size());
  const size_t K = size_t(m.K);
  x.resize(K * Np);
  y.resize(K * Np);
  for (size_t c = 0; c < K; ++c)
    for (int j = 0; j < Np; ++j) {
      const double l1 = 0.5 * (s[j] + 1.0), l2 = -0.5 * (r[j] + s[j]), l3 = 0.5 * (r[j] + 1.0);
      const size_t idx = node_major ? size_t(j) * K + c : c * Np + j;
      x[idx] = l2 * m.cx[3 * c] + l3 * m.cx[3 * c + 1] + l1 * m.cx[3 * c + 2];
      y[idx] = l2 * m.cy[3 * c] + l3 * m.cy[3 * c + 1] + l1 * m.cy[3 * c + 2];
    }
}

bool write_gmsh41(const Mesh2D& m, const std::string& path, std::string& err) {
  if (m.periodic) {
    err = "a periodic mesh has no MSH 4.1 representation";
    return false;
  }
  std::ofstream out(path);
  if (!out) {
    err = "cannot open " + path;
    return false;
  }
  out.precision(17);
  out << "$MeshFormat\n4.1 0 8\n$EndMeshFormat\n";
  // nodes: numEntityBlocks numNodes minNodeTag maxNodeTag, then one block
  // (entityDim 2, entityTag 1, parametric 0) with its tags and coordinates
  out << "$Nodes\n1 " << m.Nv << " 1 " << m.Nv << "\n2 1 0 " << m.Nv << "\n";
  for (int v = 1; v <= m.Nv; ++v) out << v << "\n";
  for (int v = 0; v < m.Nv; ++v) out << m.px[size_t(v)] << " " << m.py[size_t(v)] << " 0\n";
  out << "$EndNodes\n";
  // elements: one block of 3-node triangles (element type 2)
  out << "$Elements\n1 " << m.K << " 1 " << m.K << "\n2 1 2 " << m.K << "\n";
  for (int c = 0; c < m.K; ++c)
    out << c + 1 << " " << m.etov[3 * size_t(c)] + 1 << " " << m.etov[3 * size_t(c) + 1] + 1 << " "
        << m.etov[3 * size_t(c) + 2] + 1 << "\n";
  out << "$EndElements\n";
  if (!out) {
    err = "write failed: " + path;
    return false;
  }
  return true;
}

bool read_gmsh41(const std::string& path, std::vector<double>& vx, std::vector<double>& vy,
                 std::vector<int>& etov, std::string& err) {
  std::ifstream in(path);
  if (!in) {
    err = "cannot open " + path;
    return false;
  }
  std::string line;
  vx.clear();
  vy.clear();
  etov.clear();
  long min_tag = 1;
  while (std::getline(in, line)) {
    if (line.rfind("$MeshFormat", 0) == 0) {
      std::getline(in, line);
      if (std::stod(line) < 4.1) {
        err = "MSH version below 4.1";
        return false;
      }
    } else if (line.rfind("$Nodes", 0) == 0) {
      long nblocks, nnodes, tmin, tmax;
      in >> nblocks >> nnodes >> tmin >> tmax;
      min_tag = tmin;
      vx.assign(size_t(tmax - tmin + 1), 0.0);
      vy.assign(size_t(tmax - tmin + 1), 0.0);
      for (long b = 0; b < nblocks; ++b) {
        long edim, etag, param, n;
        in >> edim >> etag >> param >> n;
        std::vector<long> tags(n);
        for (long i = 0; i < n; ++i) in >> tags[i];
        for (long i = 0; i < n; ++i) {
          double x, y, z;
          in >> x >> y >> z;
          vx[size_t(tags[i] - tmin)] = x;
          vy[size_t(tags[i] - tmin)] = y;
        }
      }
    } else if (line.rfind("$Elements", 0) == 0) {
      long nblocks, nelem, tmin, tmax;
      in >> nblocks >> nelem >> tmin >> tmax;
      for (long b = 0; b < nblocks; ++b) {
        long edim, etag, type, n;
        in >> edim >> etag >> type >> n;
        std::getline(in, line);
        for (long i = 0; i < n; ++i) {
          std::getline(in, line);
          if (type != 2) continue;
          std::istringstream ls(line);
          long tag, a, bb, c;
          ls >> tag >> a >> bb >> c;
          etov.push_back(int(a - min_tag));
          etov.push_back(int(bb - min_tag));
          etov.push_back(int(c - min_tag));
        }
      }
    }
  }
  if (vx.empty() || etov.empty()) {
    err = "no nodes or triangles found";
    return false;
  }
  return true;
}

}  // namespace dgc_host
