// ORACLE -- test infrastructure only.
//
// Driver that runs the reference's own serial CPU path (USE_CPU_ONLY) from
// its headers where they lie under /root/reference (compiled by
// oracle/build_ref.sh into oracle/_ref/; nothing of the reference is copied).
// It replaces the example main() programs, which need the CUDA runtime
// (examples/maxwell_2d/main.cpp:30-34), with file in / file out commands used
// to write the golden vectors under tests/golden/ and to time the reference
// CPU path as the bench's cpu_baseline.
//
// Mesh file (little endian): int32 nv, int32 k, f64 vx[nv], f64 vy[nv], int32 etov[3k].
// Output files are raw f64 arrays, described per command below.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include <thrust/host_vector.h>
#include <thrust/iterator/zip_iterator.h>

#include "advection_1d.h"
#include "advection_2d.h"
#include "axpy.h"
#include "div_2d.h"
#include "grad_2d.h"
#include "explicit_runge_kutta.h"
#include "maxwell_2d.h"
#include "poisson_2d.h"
#include "simple_triangular_mesh_2d.h"

using Mesh = dgc::simple_triangular_mesh_2d<double, int>;

static void die(const std::string& m) {
  std::cerr << "ref_harness: " << m << std::endl;
  std::exit(2);
}

static Mesh load_mesh(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) die("cannot open mesh " + path);
  int32_t nv, k;
  f.read((char*)&nv, 4);
  f.read((char*)&k, 4);
  std::vector<double> vx(nv), vy(nv);
  std::vector<int32_t> etov(3 * size_t(k));
  f.read((char*)vx.data(), nv * 8);
  f.read((char*)vy.data(), nv * 8);
  f.read((char*)etov.data(), 3 * size_t(k) * 4);
  if (!f) die("short mesh file " + path);
  std::vector<double> coords;
  for (int i = 0; i < nv; ++i) {
    coords.push_back(vx[i]);
    coords.push_back(vy[i]);
  }
  Mesh m;
  m.fill_vertices(coords.begin(), coords.end());
  // src/mesh/simple_triangular_mesh_2d.h:65 evaluates its two coordinate
  // reads unsequenced; refuse a build that stores them swapped.
  for (int i = 0; i < nv; ++i) {
    const auto v = m.get_vertex(i);
    if (v.x() != vx[i] || v.y() != vy[i]) die("fill_vertices stored (x, y) swapped; build with clang/hipcc");
  }
  std::vector<int> conn(etov.begin(), etov.end());
  m.fill_connectivity(conn.begin(), conn.end());
  m.build_topology();
  return m;
}

static void write(std::ofstream& f, const std::vector<double>& v) { f.write((const char*)v.data(), v.size() * 8); }

template <class It>
static void write_it(std::ofstream& f, It b, It e) {
  for (; b != e; ++b) {
    double d = double(*b);
    f.write((const char*)&d, 8);
  }
}

static void write_mat(std::ofstream& f, const dgc::dense_matrix<double, false>& m) {
  std::vector<double> v(m.size_row() * m.size_col());
  for (std::size_t i = 0; i < m.size_row(); ++i)
    for (std::size_t j = 0; j < m.size_col(); ++j) v[i * m.size_col() + j] = m(i, j);
  write(f, v);
}

using HVec = thrust::host_vector<double>;
using HIt = HVec::iterator;
using HZip = thrust::zip_iterator<thrust::tuple<HIt, HIt, HIt>>;

struct Fields3 {
  HVec a, b, c;
  explicit Fields3(std::size_t n) : a(n), b(n), c(n) {}
  HZip zip() { return thrust::make_zip_iterator(thrust::make_tuple(a.begin(), b.begin(), c.begin())); }
  std::vector<double> flat() const {
    std::vector<double> v(a.begin(), a.end());
    v.insert(v.end(), b.begin(), b.end());
    v.insert(v.end(), c.begin(), c.end());
    return v;
  }
};

int main(int argc, char** argv) {
  if (argc < 2) die("usage: ref_harness <command> ...");
  const std::string cmd = argv[1];

  if (cmd == "refelem" && argc == 5) {
    // out: r[Np] s[Np] f0 f1 f2 [Nfp each] Dr_w Ds_w L_w (Maxwell) Dr_s Ds_s L_s (Poisson)
    Mesh m = load_mesh(argv[2]);
    const int N = std::atoi(argv[3]);
    maxwell_2d<double, Mesh> mx(m, N);
    poisson_2d<double, Mesh> po(m, N);
    std::vector<std::pair<double, double>> pos;
    dgc::reference_triangle<double>::node_positions(N, std::back_inserter(pos));
    std::ofstream f(argv[4], std::ios::binary);
    for (auto& p : pos) f.write((const char*)&p.first, 8);
    for (auto& p : pos) f.write((const char*)&p.second, 8);
    write_it(f, mx.F0_Nodes.begin(), mx.F0_Nodes.end());
    write_it(f, mx.F1_Nodes.begin(), mx.F1_Nodes.end());
    write_it(f, mx.F2_Nodes.begin(), mx.F2_Nodes.end());
    write_mat(f, mx.m_Dr);
    write_mat(f, mx.m_Ds);
    write_mat(f, mx.m_L);
    write_mat(f, po.m_Dr);
    write_mat(f, po.m_Ds);
    write_mat(f, po.m_L);
    return 0;
  }

  if (cmd == "geometry" && argc == 5) {
    // out: invJ[4K] J[K] faceJ[3K] iface_cell[3K] iface_face[3K] nx[3K] ny[3K] nb bx[nb] by[nb]
    Mesh m = load_mesh(argv[2]);
    const int N = std::atoi(argv[3]);
    maxwell_2d<double, Mesh> op(m, N);
    std::vector<double> ij, J, fJ, bx, by, nx, ny;
    std::vector<int> ic, iff;
    op.fill_cell_mappings(std::back_inserter(ij), std::back_inserter(J), std::back_inserter(fJ));
    op.fill_cell_interfaces(std::back_inserter(ic), std::back_inserter(iff));
    const int nb = op.fill_boundary_nodes(std::back_inserter(bx), std::back_inserter(by));
    op.fill_outward_normals(std::back_inserter(nx), std::back_inserter(ny));
    std::ofstream f(argv[4], std::ios::binary);
    write(f, ij);
    write(f, J);
    write(f, fJ);
    write_it(f, ic.begin(), ic.end());
    write_it(f, iff.begin(), iff.end());
    write(f, nx);
    write(f, ny);
    double dnb = nb;
    f.write((const char*)&dnb, 8);
    write(f, bx);
    write(f, by);
    return 0;
  }

  if (cmd == "maxwell" && argc == 7) {
    // Maxwell TM cavity (examples/maxwell_2d/main.cpp:65-231), CPU layout.
    // args: mesh N T dt out. out: x y IC RHS(IC) step1 final [3n each for fields]
    //       then f64 steps, f64 t_final, f64 seconds of the time loop.
    Mesh m = load_mesh(argv[2]);
    const int N = std::atoi(argv[3]);
    const double T = std::atof(argv[4]);
    double dt = std::atof(argv[5]);
    maxwell_2d<double, Mesh> op(m, N);
    const int n = op.total_num_nodes();
    thrust::host_vector<dgc::point_2d<double>> x(n);
    Fields3 u(n), w0(n), w1(n), w2(n), w3(n), w4(n), rhs(n);
    op.initialize_dofs(x.begin(), u.zip());
    std::vector<double> xs(n), ys(n);
    for (int i = 0; i < n; ++i) {
      xs[i] = x[i].x();
      ys[i] = x[i].y();
    }
    const std::vector<double> ic = u.flat();
    op(u.zip(), n, 0.0, rhs.zip());
    double t = 0.0;
    int steps = 0;
    std::vector<double> step1;
    auto t0 = std::chrono::steady_clock::now();
    while (t < T) {
      if (t + dt > T) dt = T - t;
      dgc::rk4(u.zip(), n, t, dt, op, &dgc::axpy_n<double, HZip, HZip>, w0.zip(), w1.zip(), w2.zip(),
               w3.zip(), w4.zip());
      t += dt;
      if (++steps == 1) step1 = u.flat();
    }
    auto t1 = std::chrono::steady_clock::now();
    std::ofstream f(argv[6], std::ios::binary);
    write(f, xs);
    write(f, ys);
    write(f, ic);
    write(f, rhs.flat());
    write(f, step1);
    write(f, u.flat());
    const double meta[3] = {double(steps), t, std::chrono::duration<double>(t1 - t0).count()};
    f.write((const char*)meta, sizeof meta);
    return 0;
  }

  if (cmd == "maxwell_time" && argc == 6) {
    // Timing only: args mesh N nsteps dt; prints "steps seconds" for the time loop.
    Mesh m = load_mesh(argv[2]);
    const int N = std::atoi(argv[3]);
    const int nsteps = std::atoi(argv[4]);
    const double dt = std::atof(argv[5]);
    maxwell_2d<double, Mesh> op(m, N);
    const int n = op.total_num_nodes();
    thrust::host_vector<dgc::point_2d<double>> x(n);
    Fields3 u(n), w0(n), w1(n), w2(n), w3(n), w4(n);
    op.initialize_dofs(x.begin(), u.zip());
    double t = 0.0;
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < nsteps; ++i) {
      dgc::rk4(u.zip(), n, t, dt, op, &dgc::axpy_n<double, HZip, HZip>, w0.zip(), w1.zip(), w2.zip(),
               w3.zip(), w4.zip());
      t += dt;
    }
    auto t1 = std::chrono::steady_clock::now();
    double chk = 0.0;
    for (int i = 0; i < n; ++i) chk += u.c[i];
    std::printf("%d %.9f %.17g\n", nsteps, std::chrono::duration<double>(t1 - t0).count(), chk);
    return 0;
  }

  if (cmd == "advection" && argc == 8) {
    // Reference advection problem (a = (1,1), analytic inflow), CPU layout.
    // args: mesh N t_rhs nsteps dt out. out: IC, RHS(IC, t_rhs), state after nsteps from t = 0.
    Mesh m = load_mesh(argv[2]);
    const int N = std::atoi(argv[3]);
    const double trhs = std::atof(argv[4]);
    const int nsteps = std::atoi(argv[5]);
    const double dt = std::atof(argv[6]);
    advection_2d<double, Mesh> op(m, N);
    const int n = op.total_num_nodes();
    std::vector<dgc::point_2d<double>> x(n);
    std::vector<double> u(n), rhs(n), w0(n), w1(n), w2(n), w3(n), w4(n);
    op.initialize_dofs(x.begin(), u.begin());
    const std::vector<double> ic = u;
    op(u.cbegin(), n, trhs, rhs.begin());
    double t = 0.0;
    for (int i = 0; i < nsteps; ++i) {
      dgc::rk4(u.data(), n, t, dt, op, &dgc::axpy_n<double, const double*, double*>, w0.data(), w1.data(),
               w2.data(), w3.data(), w4.data());
      t += dt;
    }
    std::ofstream f(argv[7], std::ios::binary);
    write(f, ic);
    write(f, rhs);
    write(f, u);
    return 0;
  }

  if (cmd == "advection_time" && argc == 6) {
    // Timing only: args mesh N nsteps dt; prints "steps seconds checksum" for
    // the reference advection time loop (examples/advection_2d/main.cpp:135-152).
    Mesh m = load_mesh(argv[2]);
    const int N = std::atoi(argv[3]);
    const int nsteps = std::atoi(argv[4]);
    const double dt = std::atof(argv[5]);
    advection_2d<double, Mesh> op(m, N);
    const int n = op.total_num_nodes();
    std::vector<dgc::point_2d<double>> x(n);
    std::vector<double> u(n), w0(n), w1(n), w2(n), w3(n), w4(n);
    op.initialize_dofs(x.begin(), u.begin());
    double t = 0.0;
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < nsteps; ++i) {
      dgc::rk4(u.data(), n, t, dt, op, &dgc::axpy_n<double, const double*, double*>, w0.data(), w1.data(),
               w2.data(), w3.data(), w4.data());
      t += dt;
    }
    auto t1 = std::chrono::steady_clock::now();
    double chk = 0.0;
    for (double v : u) chk += v;
    std::printf("%d %.9f %.17g\n", nsteps, std::chrono::duration<double>(t1 - t0).count(), chk);
    return 0;
  }

  if (cmd == "laplace" && argc == 6) {
    // poisson_2d::operator() (laplace_2d with homogeneous Dirichlet) applied to
    // each vector of an input file of nvec * K*Np f64 (CPU layout).
    // args: mesh N infile out. out: the results, same shape.
    Mesh m = load_mesh(argv[2]);
    const int N = std::atoi(argv[3]);
    poisson_2d<double, Mesh> op(m, N);
    const std::size_t n = op.total_num_nodes();
    std::ifstream in(argv[4], std::ios::binary | std::ios::ate);
    const std::size_t bytes = in.tellg();
    in.seekg(0);
    if (bytes % (8 * n) != 0) die("laplace: input size is not a multiple of K*Np");
    const std::size_t nvec = bytes / (8 * n);
    std::ofstream f(argv[5], std::ios::binary);
    std::vector<double> v(n);
    for (std::size_t k = 0; k < nvec; ++k) {
      in.read((char*)v.data(), 8 * n);
      op(v.begin());
      write(f, v);
    }
    return 0;
  }

  if ((cmd == "grad" || cmd == "div") && argc == 7) {
    // grad_2d::operator() (src/grad_2d.h:152-187) on each vector, or
    // div_2d::operator() (src/div_2d.h:154-188) on each pair of vectors, of an
    // input file of f64 (CPU layout, K*Np per vector). BC (exterior value at
    // boundary node (x, y), interior value a): bc 0: a; bc 1: 2 sin(x) cos(y) - a
    // (first component), 2 cos(x) sin(y) - a (second). args: mesh N bc infile out.
    // out: grad: per vector the x then the y component; div: per pair the divergence.
    Mesh m = load_mesh(argv[2]);
    const int N = std::atoi(argv[3]);
    const int bc = std::atoi(argv[4]);
    struct BC1 {
      int kind;
      double exterior_val(double x, double y, double a) const { return kind ? 2.0 * std::sin(x) * std::cos(y) - a : a; }
    };
    struct BC2 {
      int kind;
      std::pair<double, double> exterior_val(double x, double y, std::pair<double, double> a) const {
        if (!kind) return a;
        return std::make_pair(2.0 * std::sin(x) * std::cos(y) - a.first, 2.0 * std::cos(x) * std::sin(y) - a.second);
      }
    };
    const std::size_t n = std::size_t(m.num_cells()) * (N + 1) * (N + 2) / 2;
    std::ifstream in(argv[5], std::ios::binary | std::ios::ate);
    const std::size_t bytes = in.tellg();
    in.seekg(0);
    const std::size_t per = cmd == "grad" ? n : 2 * n;
    if (bytes % (8 * per) != 0) die(cmd + ": input size is not a multiple of the vector size");
    const std::size_t nvec = bytes / (8 * per);
    std::ofstream f(argv[6], std::ios::binary);
    std::vector<double> v0(n), v1(n);
    if (cmd == "grad") {
      dgc::grad_2d<double, Mesh> op(m, N);
      for (std::size_t k = 0; k < nvec; ++k) {
        in.read((char*)v0.data(), 8 * n);
        op(v0.begin(), v1.begin(), BC1{bc});
        write(f, v0);
        write(f, v1);
      }
    } else {
      dgc::div_2d<double, Mesh> op(m, N);
      for (std::size_t k = 0; k < nvec; ++k) {
        in.read((char*)v0.data(), 8 * n);
        in.read((char*)v1.data(), 8 * n);
        op(v0.begin(), v1.begin(), BC2{bc});
        write(f, v0);
      }
    }
    return 0;
  }

  if (cmd == "laplace_bc" && argc == 7) {
    // laplace_2d (src/laplace_2d.h:44-81) with a BC object of one kind on every
    // boundary node. kind 0 (Dirichlet): exterior_val(x, y, a) = 2 sin(x) cos(y) - a;
    // kind 1 (Neumann): exterior_grad_n(x, y, g) = 2 (cos(x) sin(y) + x y) - g.
    // args: mesh N kind infile out. out: apply_homogeneous of every input vector,
    // then apply_inhomogeneous_as_rhs of a zero vector, then operator() of every vector.
    Mesh m = load_mesh(argv[2]);
    const int N = std::atoi(argv[3]);
    const int kind = std::atoi(argv[4]);
    struct BC {
      int kind;
      bool is_dirichlet(double, double) const { return kind == 0; }
      double exterior_val(double x, double y, double a) const { return kind == 0 ? 2.0 * std::sin(x) * std::cos(y) - a : a; }
      double exterior_grad_n(double x, double y, double g) const {
        return kind == 0 ? g : 2.0 * (std::cos(x) * std::sin(y) + x * y) - g;
      }
    };
    const BC bc{kind};
    dgc::laplace_2d<double, Mesh> op(m, N);
    const std::size_t n = op.total_num_nodes();
    std::ifstream in(argv[5], std::ios::binary | std::ios::ate);
    const std::size_t bytes = in.tellg();
    in.seekg(0);
    if (bytes % (8 * n) != 0) die("laplace_bc: input size is not a multiple of K*Np");
    const std::size_t nvec = bytes / (8 * n);
    std::vector<std::vector<double>> vin(nvec, std::vector<double>(n));
    for (auto& v : vin) in.read((char*)v.data(), 8 * n);
    std::ofstream f(argv[6], std::ios::binary);
    for (auto v : vin) {
      op.apply_homogeneous(v.begin(), bc);
      write(f, v);
    }
    std::vector<double> z(n, 0.0);
    op.apply_inhomogeneous_as_rhs(z.begin(), bc);
    write(f, z);
    for (auto v : vin) {
      op(v.begin(), bc);
      write(f, v);
    }
    return 0;
  }

  if (cmd == "adv1d" && argc == 6) {
    // examples/advection_1d: advection_1d<double>(numCells, order) and the
    // loop of its main.cpp:105-116 (rk4 with axpy_n, dt = 0.25 / N^2 h / a).
    // args: numCells order nsteps out. out: m_D, m_L (row-major), x, the
    // initial state, its right-hand side at t = 0.3, the state after nsteps
    // (CPU layout, cell-major).
    const int K = std::atoi(argv[2]);
    const int N = std::atoi(argv[3]);
    const int nsteps = std::atoi(argv[4]);
    advection_1d<double> op(K, N);
    const int n = op.num_dofs();
    std::vector<double> x(n), v(n), r(n), w1(n), w2(n), w3(n), w4(n), w5(n);
    op.initialize_dofs(x.begin(), v.begin());
    std::vector<double> v0 = v;
    op(v0.cbegin(), n, 0.3, r.begin());
    double t = 0.0;
    const double dt = 0.25 / N / N * op.min_elem_size() / op.wave_speed();
    for (int i = 0; i < nsteps; ++i) {
      dgc::rk4(v.data(), n, t, dt, op, &dgc::axpy_n<double, const double*, double*>, w1.data(), w2.data(),
               w3.data(), w4.data(), w5.data());
      t += dt;
    }
    std::ofstream f(argv[5], std::ios::binary);
    write_mat(f, op.m_D);
    write_mat(f, op.m_L);
    write(f, x);
    write(f, v0);
    write(f, r);
    write(f, v);
    return 0;
  }

  if (cmd == "laplace_time" && argc == 5) {
    // Timing only: args mesh N napply; prints "napply seconds".
    Mesh m = load_mesh(argv[2]);
    const int N = std::atoi(argv[3]);
    const int napply = std::atoi(argv[4]);
    poisson_2d<double, Mesh> op(m, N);
    const std::size_t n = op.total_num_nodes();
    std::vector<double> v(n);
    for (std::size_t i = 0; i < n; ++i) v[i] = std::sin(0.001 * double(i));
    auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < napply; ++k) {
      op(v.begin());
      for (auto& z : v) z *= 1e-3;
    }
    auto t1 = std::chrono::steady_clock::now();
    std::printf("%d %.9f\n", napply, std::chrono::duration<double>(t1 - t0).count());
    return 0;
  }

  die("unknown command or wrong argument count: " + cmd);
}
