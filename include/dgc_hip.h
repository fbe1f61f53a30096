/*
 * dgc_hip.h -- C ABI of the MI355X (gfx950) nodal-DG hot path.
 *
 * Plain pointers and sizes only. Every entry point returns an int status
 * (DGC_OK == 0); dgc_last_error() gives the message of the last failure on
 * the calling thread. Device work is queued on the context's stream and is
 * asynchronous unless stated otherwise; handles are not thread-safe.
 *
 * Array conventions follow the reference:
 *   - DOFs on the device are node-major per field: u[j * K + c]
 *     (examples/maxwell_2d/maxwell_2d.h:129-135, the !USE_CPU_ONLY branch);
 *   - host geometry arrays are the outputs of simple_discretization_2d::fill_*
 *     (src/simple_discretization_2d.h:112-218): inv_jac[4K] = (rx, ry, sx, sy),
 *     J[K], face_J[3K], iface_cell/iface_face[3K] (boundary face: cell itself and
 *     the offset into the boundary-node arrays), nx/ny[3K];
 *   - operator matrices are row-major (src/dense_matrix.h, CM = false).
 */
#ifndef DGC_HIP_H
#define DGC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum dgc_status {
  DGC_OK = 0,
  DGC_ERR_INVALID = 1,       /* bad argument (size, order, null pointer) */
  DGC_ERR_HIP = 2,           /* a HIP runtime call failed */
  DGC_ERR_LAUNCH = 3,        /* kernel launch failed */
  DGC_ERR_NONFINITE = 4,     /* NaN/Inf detected in a checked result */
  DGC_ERR_NOT_CONVERGED = 5, /* iterative solver hit its iteration cap */
  DGC_ERR_UNSUPPORTED = 6    /* configuration not built into this library */
};

typedef struct dgc_ctx_s* dgc_ctx_t;
typedef struct dgc_disc_s* dgc_disc_t;
typedef struct dgc_graddiv_s* dgc_graddiv_t;
typedef struct dgc_stokes_s* dgc_stokes_t;
typedef struct dgc_adv1d_s* dgc_adv1d_t;
typedef struct dgc_op_s* dgc_op_t;
typedef struct dgc_mesh_s* dgc_mesh_t;

/* ---- errors / version -------------------------------------------------- */
const char* dgc_last_error(void);
int dgc_max_order(void);

/* ---- context and device memory ------------------------------------------
 * Replaces the legacy default stream and the thrust::device_vector /
 * cudaMallocManaged buffers of examples/maxwell_2d/main.cpp:103-114,177-197. */
int dgc_ctx_create(int device, dgc_ctx_t* out);
int dgc_ctx_destroy(dgc_ctx_t ctx);
int dgc_ctx_sync(dgc_ctx_t ctx);
int dgc_malloc(dgc_ctx_t ctx, size_t bytes, void** ptr);
int dgc_free(dgc_ctx_t ctx, void* ptr);
int dgc_memcpy_h2d(dgc_ctx_t ctx, void* dst, const void* src, size_t bytes);
int dgc_memcpy_d2h(dgc_ctx_t ctx, void* dst, const void* src, size_t bytes);
int dgc_memcpy_d2d(dgc_ctx_t ctx, void* dst, const void* src, size_t bytes);
int dgc_memset_zero(dgc_ctx_t ctx, void* dst, size_t bytes);

/* ---- host set-up (reference element, mesh, geometry) --------------------
 * dgc_ref_operators: nodes, face node lists and the operator matrices of
 * order `order` in weak form (form 0; examples/maxwell_2d/maxwell_2d.h:79-112,
 * examples/advection_2d/advection_2d.h:74-101) or strong form (form 1;
 * src/laplace_2d.h:87-112). Sizes: r, s [Np]; fmask [3 Nfp]; Dr, Ds, mass
 * [Np*Np]; lift [Np*3Nfp]. Any output pointer may be NULL. */
int dgc_ref_operators(int order, int form, double* r, double* s, int* fmask, double* Dr, double* Ds,
                      double* lift, double* mass);

/* Structured jittered triangle mesh (the generated meshes of the benchmark
 * configurations); topology as src/mesh/simple_triangular_mesh_2d.h:151-217. */
int dgc_mesh_structured(int nx, int ny, double x0, double x1, double y0, double y1, double jitter,
                        uint64_t seed, int periodic, dgc_mesh_t* out);
/* Mesh from vertex coordinates and triangle connectivity
 * (src/mesh/simple_triangular_mesh_2d.h:60-80; vertices kept as (x, y)). */
int dgc_mesh_from_arrays(int nv, const double* vx, const double* vy, int k, const int* etov,
                         dgc_mesh_t* out);
/* gmsh MSH 4.1 triangle mesh (src/mesh/gmsh_importer_exporter.h:49-173). */
int dgc_mesh_read_gmsh(const char* path, dgc_mesh_t* out);
/* Write a mesh as gmsh MSH 4.1 ASCII, one node and one triangle block
 * (src/mesh/gmsh_importer_exporter.h:175-213); periodic meshes are rejected. */
int dgc_mesh_write_gmsh(dgc_mesh_t mesh, const char* path);
int dgc_mesh_destroy(dgc_mesh_t mesh);
int dgc_mesh_info(dgc_mesh_t mesh, int* num_cells, int* num_vertices, double* min_edge);
/* Per-cell vertex coordinates cx, cy [3K], vertex ids etov [3K] and the twin
 * half-edge map nb_cell, nb_face [3K] (src/mesh/simple_triangular_mesh_2d.h:116-121). */
int dgc_mesh_cells(dgc_mesh_t mesh, double* cx, double* cy, int* etov, int* nb_cell, int* nb_face);
/* Geometry of order `order` (src/simple_discretization_2d.h:84-218). Call once
 * with bx == by == NULL to learn *num_boundary_nodes. */
int dgc_mesh_geometry(dgc_mesh_t mesh, int order, double* inv_jac, double* J, double* face_J,
                      double* nx, double* ny, int* iface_cell, int* iface_face,
                      int* num_boundary_nodes, double* bx, double* by);
/* Physical DOF coordinates, node-major (node_major = 1) or cell-major. */
int dgc_mesh_dof_coords(dgc_mesh_t mesh, int order, int node_major, double* x, double* y);

/* ---- device discretization ----------------------------------------------
 * Replaces create_simple_discretization_2d_on_device /
 * destroy_simple_discretization_2d_on_device (src/d_simple_discretization_2d.cuh:89-152):
 * same argument list, but the library owns the device copies and returns a
 * handle. Orders 1..dgc_max_order() are accepted (the reference: 1..6). */
int dgc_disc2d_create(dgc_ctx_t ctx, int num_cells, int order, const int* face0_nodes,
                      const int* face1_nodes, const int* face2_nodes, const double* inv_jac,
                      const double* J, const double* face_J, const int* iface_cell,
                      const int* iface_face, int num_boundary_nodes, const double* bx,
                      const double* by, const double* nx, const double* ny, dgc_disc_t* out);
int dgc_disc2d_destroy(dgc_disc_t disc);

/* Device set-up: the same discretization built on the GPU from the cells
 * alone, for meshes where the host set-up dominates (K >= 1e6). Replaces the
 * host marshalling of src/simple_discretization_2d.h:84-218 (geometric factors,
 * src/mapping_2d.h:101-158; outward normals, src/basic_geom_2d.h:110-153) and
 * the twin half-edge search of src/mesh/simple_triangular_mesh_2d.h:151-217.
 * cx, cy: [3 K] host arrays, the counter-clockwise corner coordinates of each
 * cell (periodic meshes: unwrapped); etov: [3 K] host array of vertex ids in
 * [0, num_vertices), equal ids meaning a shared vertex. Every device and host
 * array of the result equals, bit for bit, what dgc_mesh_geometry +
 * dgc_disc2d_create give for the same mesh (boundary nodes numbered in cell,
 * then face order). Errors: DGC_ERR_INVALID for a vertex id out of range, a
 * degenerate or non-manifold edge, or a clockwise / degenerate cell. */
int dgc_disc2d_create_from_cells(dgc_ctx_t ctx, int num_cells, int num_vertices, int order,
                                 const double* cx, const double* cy, const int* etov,
                                 dgc_disc_t* out);
/* Copies of a discretization's device arrays into host memory (any pointer may
 * be NULL): geo [13][K] (rx, ry, sx, sy, face_J/J for faces 0-2, nx for faces
 * 0-2, ny for faces 0-2), nbr [3][K] (twin cell << 2 | twin face, or boundary
 * node offset << 2 | 3), J [K], face_J [3][K], and the boundary face nodes.
 * No counterpart in the reference (its device copies are plain pointers the
 * caller can read). */
int dgc_disc2d_arrays(dgc_disc_t disc, double* geo, int* nbr, double* J, double* face_J,
                      int* num_boundary_nodes, double* bx, double* by);

/* ---- gradient and divergence with central fluxes -----------------------
 * Replace grad_2d (src/grad_2d.h:36-187) and div_2d (src/div_2d.h:36-188),
 * the operators either side of the Poisson solve in the Stokes splitting.
 * Dr, Ds: weak-form derivative matrices (V Vr^T M), lift: LIFT, all
 * row-major as dgc_ref_operators(order, 0, ...) returns them. The reference's
 * BC object becomes its affine form: on boundary node b (numbered cell by
 * cell, face by face, as dgc_disc2d_arrays' bx, by) the exterior value is
 * g[b] + beta * (interior value); g == NULL means g = 0 (beta = 1: the
 * "exterior = interior" BC of stokes_2d.h:162-163; g = 2 u_D, beta = -1: a
 * Dirichlet value u_D). u, g, out: device arrays, node-major [Np][K]; out of
 * place (in place in the reference: the caller copies first). Asynchronous
 * on the context's stream. */
int dgc_graddiv_create(dgc_disc_t disc, const double* Dr, const double* Ds, const double* lift,
                       dgc_graddiv_t* out);
int dgc_graddiv_destroy(dgc_graddiv_t op);
/* Waves per 64-cell tile of the kernels: 1, 2 or 4 (4 / wt tiles per
 * 256-thread workgroup; a tiling whose LDS would exceed 64 KB runs as 4), or 0
 * for the default of the order (1 at N <= 2, 2 at N = 3, 5, else 4). The
 * tilings agree to rounding (the compiler may contract the metric products
 * differently in each). */
int dgc_graddiv_set_tiling(dgc_graddiv_t op, int waves_per_tile);
/* (out_x, out_y) = grad u  (grad_2d::operator(), src/grad_2d.h:152-187) */
int dgc_grad2d_apply(dgc_graddiv_t op, const double* u, const double* g, double beta, double* out_x,
                     double* out_y);
/* out = div (u0, u1)  (div_2d::operator(), src/div_2d.h:154-188); g0, g1: the
 * boundary data of the two components */
int dgc_div2d_apply(dgc_graddiv_t op, const double* u0, const double* u1, const double* g0,
                    const double* g1, double beta, double* out);

/* ---- Stokes dual-splitting step (examples/stokes_2d) ---------------------
 * stokes_2d::advance_timestep (examples/stokes_2d/stokes_2d.h:358-386) on the
 * GPU, the path the reference leaves as a TODO (k_stokes_2d.cu:96-106): from
 * (u^{n-1}, t_prev) and (u^n, t) to u^{n+1} at t + dt for the example's
 * analytic problem (a = 2.883356; its boundary data, stokes_2d.h:100-183),
 * with tau = 1 as the reference. Pressure: PCG on -M Lap_h with Neumann
 * faces (block-Jacobi, or two-level after dgc_stokes_coarse_space),
 * right-hand side projected onto the complement of the constants (the
 * reference factorises the singular matrix directly); viscous step:
 * block-Jacobi PCG on -M Lap_h + (3 / 2dt) M with Dirichlet faces, one solve
 * per component. rtol: relative residual of every solve.
 * Retained state: the handle keeps the last pressure, and each pressure
 * solve starts from it (zero before the first call); a new trajectory starts
 * from a fresh handle or after dgc_stokes_reset. The viscous solves start
 * from 2 u^n - u^{n-1}, written into ux_out / uy_out before the solves:
 * ux_out may alias ux or ux_prev (and uy_out uy or uy_prev), but an output
 * must not alias an input of the other component.
 * Velocities: device arrays, node-major [Np][K]. iters (may be NULL): the
 * pressure, x and y iteration counts. Synchronous. */
int dgc_stokes_create(dgc_disc_t disc, dgc_stokes_t* out);
/* Two-level pressure solves: the coarse space of dgc_ipdg_coarse_space on the
 * pressure operator (cell corners cx, cy [3K] of the mesh, gx x gy boxes). */
int dgc_stokes_coarse_space(dgc_stokes_t s, const double* cell_x, const double* cell_y, int gx, int gy);
int dgc_stokes_destroy(dgc_stokes_t s);
/* Forgets the retained pressure (the next pressure solve starts from zero). */
int dgc_stokes_reset(dgc_stokes_t s);
int dgc_stokes_advance(dgc_stokes_t s, const double* ux_prev, const double* uy_prev, double t_prev,
                       const double* ux, const double* uy, double t, double dt, double rtol,
                       double* ux_out, double* uy_out, int* iters);

/* ---- 1D advection (examples/advection_1d) --------------------------------
 * advection_1d (examples/advection_1d/advection_1d.h:38-199) and its device
 * object (d_advection_1d.cuh:37-98): u_t + 2 pi u_x = 0 on [0, 2 pi] with
 * num_cells uniform cells, inflow -sin(2 pi t), upwind flux.
 * dgc_adv1d_operators: the strong-form D = Dr 2/h and L = V V^T 2/h
 * ((order+1)^2, row-major, the host class's m_D and m_L) and the node
 * positions x [num_cells (order+1)] node-major; any output may be NULL.
 * dgc_adv1d_create takes D and L (e.g. the reference's m_D, m_L) or builds
 * them when either is NULL; orders 1..12. */
int dgc_adv1d_operators(int num_cells, int order, double* D, double* L, double* x);
int dgc_adv1d_create(dgc_ctx_t ctx, int num_cells, int order, const double* D, const double* L,
                     dgc_adv1d_t* out);
int dgc_adv1d_destroy(dgc_adv1d_t op);
/* out = the right-hand side at time t (advection_1d::operator()); node-major, out != in */
int dgc_adv1d_rhs(dgc_adv1d_t op, const double* in, double t, double* out);
/* nsteps classical RK4 steps of dt from t0, in place (the loop of
 * examples/advection_1d/main.cpp:105-116 with rk4_on_device), fused stages */
int dgc_adv1d_rk4_steps(dgc_adv1d_t op, double* u, double t0, double dt, int nsteps);

/* ---- plot data files -------------------------------------------------------
 * The text files the examples write for plotting (examples/advection_2d/
 * main.cpp:164-174 "Advection2DDataFile.txt", examples/advection_1d/
 * main.cpp:124-134 "Advection1DDataFile.txt"): 15 significant digits, the
 * solution block then the reference-solution block, with the examples'
 * headers. y == NULL writes the 1D layout. Host arrays of n values. */
int dgc_write_plot_file(const char* path, size_t n, const double* x, const double* y, const double* u,
                        const double* ref);

/* ---- semi-discrete operators (device objects) ----------------------------
 * Replace create_device_object / destroy_device_object of
 * examples/maxwell_2d/k_maxwell_2d.cu:45-116 and
 * examples/advection_2d/k_advection_2d.cu:45-126. Dr, Ds [Np*Np] and lift
 * [Np*3Nfp] are the weak-form matrices; they are copied to device memory
 * (no per-process __constant__ symbols, so several orders can coexist). */
int dgc_maxwell_tm_create(dgc_disc_t disc, const double* Dr, const double* Ds, const double* lift,
                          dgc_op_t* out);
/* a = (ax, ay); bc 0: analytic inflow sin(pi(x+1-t)) sin(pi(y+1-t))
 * (examples/advection_2d/d_advection_2d.cuh:86-91), bc 1: zero inflow.
 * ax = ay = 1 with bc 0 is the reference problem. */
int dgc_advection_create(dgc_disc_t disc, const double* Dr, const double* Ds, const double* lift,
                         double ax, double ay, int bc, dgc_op_t* out);
int dgc_op_destroy(dgc_op_t op);
int dgc_op_num_fields(dgc_op_t op);

/* Tiling of the stage kernels: wt waves per 64-cell tile (1, 2 or 4: 4/wt
 * tiles per 256-thread workgroup; 8 or 16: one tile per 512- or 1024-thread
 * workgroup, for meshes of a few tiles; 32: the output rows of each tile split
 * over 4 workgroups of 4 waves, each of which loads the tile and forms its
 * fluxes; the reference's launch-shape arguments, main.cpp argv[2], play this
 * role), or wt = 0 for the FP64
 * matrix-core variant (Dr u, Ds u and LIFT flux as 16x16x4 f64 MFMA
 * products, one 64-cell tile per workgroup; FP64 throughout). A tiling whose
 * LDS would not fit falls back to 4. dgc_op_autotune times an RK4 step (the 4
 * stage kernels) of every built tiling on the caller's buffers (work vectors only are written: stage 4
 * updates wk2 in place of the state) and keeps the fastest; ms_out[4]
 * receives the mean stage-2 kernel time of wt = 1, 2, 4, 0 (-1 where not
 * built). No reference counterpart. */
int dgc_op_set_waves_per_tile(dgc_op_t op, int wt);
/* Mean duration (*ms) of reps back-to-back launches of stage `stage` (1..4) at
 * the current tiling, HIP events on the context stream, with the buffers of
 * dgc_op_autotune (work vectors only are written). Measurement entry point;
 * no reference counterpart. */
int dgc_op_stage_time(dgc_op_t op, double* const* inout, size_t size, double* const* wk, int stage, int reps,
                      float* ms);
int dgc_op_waves_per_tile(dgc_op_t op);
/* Arithmetic of the stage contractions (no reference counterpart).
 * 0: FP64 throughout (default). 1: Dr u, Ds u and LIFT flux on the matrix
 * cores in FP32 (16x16x4 f32 MFMA, operands rounded to FP32), with the state,
 * stage accumulation, metric terms, fluxes and the stage update in FP64.
 * Both operators (Maxwell TM and advection); an ensemble takes FP64 operators. */
int dgc_op_set_precision(dgc_op_t op, int precision);
int dgc_op_precision(dgc_op_t op);
/* Cache policy of the RK stage kernels (no reference counterpart). The RK4
 * accumulator, k3 and the step's initial state are read or written once per
 * stage; with the streaming policy those accesses carry the nontemporal
 * hint, so they do not displace the stage input, whose node values the
 * neighbouring cells read again. DGC_CACHE_AUTO (default) streams when a
 * step's working set (state + 4 work vectors + 44 bytes of geometry per
 * cell) exceeds MI355X's 256 MB last-level cache. Results are identical
 * under every policy. dgc_op_cache_streaming: 1 if the streaming kernels
 * run, 0 if not, -1 for a null operator. */
enum { DGC_CACHE_AUTO = 0, DGC_CACHE_TEMPORAL = 1, DGC_CACHE_STREAMING = 2 };
int dgc_op_set_cache_policy(dgc_op_t op, int policy);
int dgc_op_cache_streaming(dgc_op_t op);
int dgc_op_autotune(dgc_op_t op, double* const* inout, size_t size, double* const* wk, int reps,
                    float* ms_out);

/* out = L_h(in, t): one evaluation of the semi-discrete operator.
 * Replaces device_SemiDiscOp_wrapper::operator() (src/device_SemiDiscOp_wrapper.cuh:45-58).
 * in/out: arrays of F device pointers (F = 3 for Maxwell: Hx, Hy, Ez), each K*Np. */
int dgc_op_rhs(dgc_op_t op, const double* const* in, size_t size, double t, double* const* out);

/* out = a * x + y over n doubles (src/k_axpy.cuh:38-59). */
int dgc_axpy(dgc_ctx_t ctx, double a, const double* x, size_t n, const double* y, double* out);

/* One classical RK4 step, fused: 4 kernels, 17 state passes.
 * Replaces rk4_on_device (examples/maxwell_2d/k_maxwell_2d.cu:98-111) driving
 * dgc::rk4 (src/explicit_runge_kutta.h:33-51). inout: F device pointers;
 * wk: 5*F device pointers (wk[w*F + f], w = 0..4), caller-owned like wk0..wk4. */
int dgc_rk4_step(dgc_op_t op, double* const* inout, size_t size, double t, double dt,
                 double* const* wk);

/* dgc_rk4_step without allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc). The 4*F work vectors live
 * in a static 64 MiB device array of the library, one per GPU, so
 * 4 * F * K * Np <= 8388608 doubles (F * K * Np <= 2097152; larger sizes
 * return DGC_ERR_INVALID). Allocates and synchronises nothing; asynchronous
 * on the context's stream. Calls that overlap on one GPU (several streams,
 * contexts or threads) share the array and race. */
int dgc_rk4_step_noalloc(dgc_op_t op, double* const* inout, size_t size, double t, double dt);

/* Time loop t0 -> t_end with step dt, the last step truncated to land on
 * t_end (examples/maxwell_2d/main.cpp:199-209). Full steps are replayed from a
 * captured hipGraph. *t_out / *steps_out report the final time and the step
 * count (either may be NULL). Blocks until the steps are done, then scans the
 * state once: DGC_ERR_NONFINITE if it holds a NaN or an infinity (the
 * reference's has_nan helper, examples/advection_2d/main.cpp:53-58). */
int dgc_rk4_run(dgc_op_t op, double* const* inout, size_t size, double t0, double t_end, double dt,
                double* const* wk, double* t_out, int* steps_out);

/* Fixed number of full steps (examples/advection_2d/main.cpp:135-152),
 * graph-replayed; blocks and checks the state as dgc_rk4_run does. */
int dgc_rk4_steps(dgc_op_t op, double* const* inout, size_t size, double t0, double dt, int nsteps,
                  double* const* wk);

/* ---- IPDG Poisson operator and solvers -----------------------------------
 * dgc_ipdg_create replaces the laplace_2d constructor (src/laplace_2d.h:83-119)
 * and the device object of examples/poisson_2d/k_poisson_2d.cu:44-97. Dr, Ds
 * [Np*Np] and lift [Np*3Nfp] are the strong-form matrices, mass [Np*Np] is
 * (V V^T)^{-1}. tau_mode 0: the reference's tau = 1 (src/laplace_2d.h:227-245);
 * tau_mode 1: tau = (N+1)(N+2)/h per face, h = min over both sides of
 * 2J/face_J. tau_scale multiplies either. Boundary faces are Dirichlet
 * (examples/poisson_2d/poisson_2d.h:73-79) unless dgc_ipdg_set_boundary says
 * Neumann. */
typedef struct dgc_ipdg_s* dgc_ipdg_t;
int dgc_ipdg_create(dgc_disc_t disc, const double* Dr, const double* Ds, const double* lift,
                    const double* mass, int tau_mode, double tau_scale, dgc_ipdg_t* out);
int dgc_ipdg_destroy(dgc_ipdg_t s);
/* In place u <- Lap_h u: laplace_2d::apply_homogeneous (src/laplace_2d.h:130-251;
 * with Dirichlet faces and zero data this is laplace_2d::operator(),
 * src/laplace_2d.h:121-128); replaces op_on_device
 * (examples/poisson_2d/k_poisson_2d.cu:99-108). u: K*Np node-major. */
int dgc_ipdg_apply(dgc_ipdg_t s, double* u);
/* One phase of the in-place apply, as launched by kernel_adapter::execute_phase
 * (src/kernel_adapter.cuh:77-82; examples/poisson_2d/k_poisson_2d.cu:105-107):
 * phase 1 computes q = grad u and the jumps into the solver's workspace,
 * phase 2 overwrites u with div q minus the lifted penalty terms. */
int dgc_ipdg_phase(dgc_ipdg_t s, int phase, double* u);
/* out = -M Lap_h u, the symmetric positive definite form (M = J * mass per cell). */
int dgc_ipdg_apply_spd(dgc_ipdg_t s, const double* u, double* out);
/* out = M f (mass-weighted right-hand side of -Lap u = f). */
int dgc_ipdg_mass(dgc_ipdg_t s, const double* f, double* out);
/* Boundary kind of every boundary face, the BC object's is_dirichlet() of
 * src/laplace_2d.h:175,218 (default DGC_BC_DIRICHLET): Dirichlet faces take
 * u+ = -u-, q+ = q- in the homogeneous apply; Neumann faces u+ = u-,
 * q+.n = -q-.n (no penalty). Switching drops the block-Jacobi inverses and
 * the coarse space (rebuilt on demand). With Neumann faces everywhere the
 * operator is singular (constants): the solvers need a right-hand side
 * orthogonal to the constants; the coarse space then inverts its Galerkin
 * matrix plus a rank-one term along the coarse constants. */
enum { DGC_BC_DIRICHLET = 0, DGC_BC_NEUMANN = 1 };
int dgc_ipdg_set_boundary(dgc_ipdg_t s, int kind);
/* Shift of the SPD form: dgc_ipdg_apply_spd, the block-Jacobi blocks and the
 * PCG solvers then use -M Lap_h + sigma M (sigma >= 0; 0 restores the plain
 * form), the Helmholtz operator of the Stokes viscous step, which the
 * reference assembles as Lap_h - (1.5/dt) I by unit vectors
 * (examples/stokes_2d/stokes_2d.h:373-380,443-460). dgc_ipdg_apply is not
 * shifted. A change drops the block inverses and the coarse space. */
int dgc_ipdg_set_shift(dgc_ipdg_t s, double sigma);
/* inout += the inhomogeneous boundary part of the Laplacian
 * (laplace_2d::apply_inhomogeneous_as_rhs, src/laplace_2d.h:253-330) for
 * boundary data per boundary node (device arrays in dgc_disc2d_arrays' bx
 * order; NULL = 0): gu = the BC object's exterior_val(x, y, 0) on Dirichlet
 * faces, hq = exterior_grad_n(x, y, 0) on Neumann faces. The full operator
 * laplace_2d::operator()(u, bc) is dgc_ipdg_apply(u) followed by subtracting
 * this part (src/laplace_2d.h:121-128). */
int dgc_ipdg_apply_inhomogeneous(dgc_ipdg_t s, const double* gu, const double* hq, double* inout);
/* In place u <- Lap_h u with boundary data, the whole BR1 operator: Dirichlet
 * faces take u+ = gd - u- (gd = exterior_val(x, y, 0)), Neumann faces
 * q+.n = hd - q-.n (hd = exterior_grad_n(x, y, 0)); device arrays per boundary
 * node, NULL = 0 (then it is dgc_ipdg_apply). Its affine part (the call on
 * u = 0) also reaches the interior faces of boundary cells, which
 * dgc_ipdg_apply_inhomogeneous leaves out as the reference does: with
 * inhomogeneous Dirichlet data only this form is consistent (the reference's
 * operator()(u, bc) error grows as the mesh is refined). No counterpart in the
 * reference; the Stokes step uses it. */
int dgc_ipdg_apply_data(dgc_ipdg_t s, const double* gd, const double* hd, double* u);
/* The operator of dgc_ipdg_apply (current boundary kind) as a sparse matrix in
 * CSR form, rows and columns in the node-major DOF order j K + c: the
 * "implicit" laplace_2d of examples/stokes_2d/stokes_2d.h:74-76, which the
 * reference assembles by one application per unit vector (stokes_2d.h:388-422);
 * here by probing a colouring of the cells (tens of applications). Call once
 * with col_idx == values == NULL for *nnz (and row_ptr [K Np + 1] if given),
 * then with host arrays row_ptr, col_idx [nnz], values [nnz]. Row block c
 * holds the columns of the cells within two faces of c (sorted), node by node. */
int dgc_ipdg_assemble(dgc_ipdg_t s, long long* nnz, long long* row_ptr, int* col_idx, double* values);
/* Launch mode of the solvers' iterations (no reference counterpart): 1
 * (default) replays captured graphs of 16 (PCG) or 8 (flexible CG)
 * iterations; 0 launches the same kernels one by one, in the same order, for
 * kernel tracers that do not follow graph launches. Results are identical. */
int dgc_ipdg_set_graphs(dgc_ipdg_t s, int enabled);
/* Builds the block-Jacobi preconditioner: the exact diagonal blocks of
 * -M Lap_h, probed with a distance-2 colouring of the cells, inverted per cell. */
int dgc_ipdg_block_jacobi(dgc_ipdg_t s, int* num_colors);
/* Copies the diagonal blocks [Np*Np][K] (entry (j,k) of cell c at (j*Np+k)*K+c). */
int dgc_ipdg_diag_blocks(dgc_ipdg_t s, double* blocks);
/* Coarse space of the two-level preconditioner (no reference counterpart):
 * cells are grouped by the box of a gx x gy grid over the mesh extent that
 * holds their centroid; each non-empty group carries the linear functions
 * 1, (x - x_a)/h, (y - y_a)/h (cell_x, cell_y: host vertex coordinates
 * [3K], the order the discretization's geometry was built from). The
 * Galerkin matrix A_c = Phi^T S Phi is formed with the operator itself and
 * inverted on the device; the inverse is kept in FP32. At most 2048 groups.
 * *num_coarse = 3 x groups. */
int dgc_ipdg_coarse_space(dgc_ipdg_t s, const double* cell_x, const double* cell_y, int gx, int gy,
                          int* num_coarse);
/* Host copies (each may be NULL): A_c [nc*nc] (FP64), its FP32 inverse
 * [nc*nc], the group of each cell [K], the basis coefficients [6][K] and the
 * reference node coordinates (r, s) [2][Np]. */
int dgc_ipdg_coarse_info(dgc_ipdg_t s, double* A, float* Ainv, int* cell_agg, float* coef, double* rs);
/* Conjugate gradients on (-M Lap_h) x = b to ||b - S x|| <= rtol ||b|| (FP64
 * true residual, restarted if the recurrence drifts). precond 0: none,
 * 1: block-Jacobi, 2: two-level (block-Jacobi + coarse correction,
 * z = P_BJ r + Phi A_c^{-1} Phi^T r). x: device output (initial guess 0).
 * history (host, may be NULL): maxit+1 squared recurrence residuals.
 * *ms: solve time (HIP events).
 * Replaces the assembly + sparse direct solve of examples/poisson_2d/main.cpp:132-223. */
int dgc_ipdg_pcg(dgc_ipdg_t s, const double* b, double* x, double rtol, int maxit, int precond,
                 int* iters, double* relres, double* history, float* ms);
/* Mixed precision: FP32 operator, vectors and block-Jacobi PCG to inner_rtol
 * inside an FP64 iterative-refinement loop whose residual b - S x is
 * recomputed in FP64 each outer step; stops at FP64 relative residual rtol. */
int dgc_ipdg_pcg_mixed(dgc_ipdg_t s, const double* b, double* x, double rtol, double inner_rtol,
                       int max_outer, int max_inner, int* outer_iters, int* inner_iters,
                       double* relres, float* ms);

/* ---- learned preconditioner (no reference counterpart) --------------------
 * A 3-layer message-passing network over the cells' face adjacency (hidden
 * width 64, bf16 weights/activations, fp32 accumulation; matrix-core forward
 * layers) maps (r, P_BJ r, cell metrics) to a correction of the block-Jacobi
 * step: z = P_BJ r + s_z c_t NN(r / s_r, P_BJ r / s_z, g). It is trained on
 * the GPU on residuals r = S e made with the solver's own operator from
 * seeded random error fields e, with the energy-norm loss
 * |z - e|_S^2 / |e|_S^2 (the error CG minimises), and used inside flexible CG. */
typedef struct dgc_mpnn_s* dgc_mpnn_t;
/* base 1: zb = P_BJ r; base 2: zb = the two-level step (needs
 * dgc_ipdg_coarse_space first; the coarse space must outlive the network). */
int dgc_mpnn_create(dgc_ipdg_t s, uint64_t seed, int base, dgc_mpnn_t* out);
int dgc_mpnn_destroy(dgc_mpnn_t m);
int dgc_mpnn_param_count(dgc_mpnn_t m);
/* width D0 of the input feature vector per cell (2 Np + 4 rounded up to 8) */
int dgc_mpnn_input_width(dgc_mpnn_t m);
/* parameters as one fp32 vector: A1 B1 b1 A2 B2 b2 A3 B3 b3 Wo bo (row-major) */
int dgc_mpnn_get_params(dgc_mpnn_t m, float* w);
int dgc_mpnn_set_params(dgc_mpnn_t m, const float* w);
/* y [K][Np] = network(X [K][D0]) (X rounded to bf16 on upload) */
int dgc_mpnn_eval(dgc_mpnn_t m, const float* X, float* y);
/* mean-square loss against T [K][Np] and its gradient [param_count] */
int dgc_mpnn_loss_grad(dgc_mpnn_t m, const float* X, const float* T, double* loss, float* grad);
/* Training pairs: for sample i the error e is a sum of modes^2 Dirichlet sine
 * modes of [-1,1]^2 with N(0,1) (m^2+n^2)^-1 amplitudes (seed seed0 + i),
 * evaluated at the DOF coordinates dof_x/dof_y (host, node-major), r = S e.
 * Also sets the target scale c_t = RMS((e - P_BJ r) / s_z) of sample 0. */
int dgc_mpnn_build_dataset(dgc_mpnn_t m, const double* dof_x, const double* dof_y, int nsamples,
                           uint64_t seed0, int modes);
/* Training pairs from the solver itself: for right-hand side i, b = M f with
 * f the Gaussian random field of dgc_mpnn_build_dataset (modes^2 sine modes,
 * N(0,1) (m^2+n^2)^-1 amplitudes, seed seed0 + i), PCG with the network's
 * base preconditioner runs to 1e-11 while its iterate x_j is recorded every
 * 16 iterations (x_0 = 0, up to `snapshots` per solve); each recorded iterate
 * gives the pair e = x* - x_j, r = S e, the residual a preconditioner meets
 * at that point of a solve. *npairs: the number of pairs (nrhs * snapshots
 * unless a solve converges within 16 (snapshots - 1) iterations). Sets c_t
 * from the first pair. */
int dgc_mpnn_build_dataset_solver(dgc_mpnn_t m, const double* dof_x, const double* dof_y, int nrhs,
                                  uint64_t seed0, int modes, int snapshots, int* npairs);
int dgc_mpnn_target_scale(dgc_mpnn_t m, double* ct);
int dgc_mpnn_set_target_scale(dgc_mpnn_t m, double ct);
/* Energy-norm loss of one residual r = S e (device, node-major) with
 * er = e.r: z = P_BJ r + s_z c_t y, loss = (z.Sz - 2 z.r + er) / er, its
 * gradient [param_count] (host, may be null) and z (device, may be null). */
int dgc_mpnn_energy_loss_grad(dgc_mpnn_t m, const double* r, double er, double* loss, float* grad, double* z);
/* Adam with cosine decay over samples [first, first+count), one sample per
 * step on the energy-norm loss; epoch_loss[epochs] receives the mean training
 * loss of each epoch. */
int dgc_mpnn_train(dgc_mpnn_t m, int first, int count, int epochs, float lr, uint64_t shuffle_seed,
                   double* epoch_loss);
int dgc_mpnn_eval_loss(dgc_mpnn_t m, int first, int count, double* loss);
/* Flexible CG on (-M Lap_h) x = b with the network preconditioner, to the FP64
 * true relative residual rtol. If the squared residual fails to drop below
 * stall_factor^2 times its value stall_window iterations earlier (or turns
 * non-finite), the solve continues with PCG on the network's base
 * preconditioner (block-Jacobi or two-level) from the current iterate
 * (*fell_back = 1; 2 = finished by a PCG restart after drift). A network
 * belongs to the operator it was created for: after dgc_ipdg_set_boundary or
 * dgc_ipdg_set_shift this call and the training / evaluation entry points
 * return DGC_ERR_INVALID (create and train a new network). */
int dgc_ipdg_fcg(dgc_ipdg_t s, dgc_mpnn_t m, const double* b, double* x, double rtol, int maxit,
                 int stall_window, double stall_factor, int* iters, double* relres, int* fell_back,
                 double* history, float* ms);

/* Device self-test of the matrix-core operand layouts the kernels rely on:
 * kind 0: C[32x32] = A[32x8] B[8x32] with bf16 inputs (rounded on the device),
 * kind 1: C[16x16] = A[16x4] B[4x16] in fp32, kind 2: the same in fp64
 * (inputs widened, result rounded to fp32; C holds the raw result registers),
 * kind 3: C[32x32] = A[32x16] B[16x32] with bf16 inputs (gfx950's 32x32x16
 * instruction). Row-major host arrays; the result is scattered with the
 * documented lane layout. No reference counterpart. */
int dgc_selftest_mfma(dgc_ctx_t ctx, int kind, const float* A, const float* B, float* C);

/* Same as dgc_rk4_steps, bracketed by HIP events on the context stream;
 * blocks until done and returns the elapsed milliseconds in *ms (the state
 * check runs after the timed region; measurement entry point, no reference
 * counterpart). */
int dgc_rk4_steps_timed(dgc_op_t op, double* const* inout, size_t size, double t0, double dt,
                        int nsteps, double* const* wk, float* ms);

/* Runs nsteps single steps with each of the 4 stage kernels bracketed by HIP
 * events; stage_ms[4] receives the mean duration of each stage kernel
 * (measurement entry point; no reference counterpart). */
int dgc_rk4_profile_stages(dgc_op_t op, double* const* inout, size_t size, double t0, double dt,
                           int nsteps, double* const* wk, float* stage_ms);

/* ---- batched ensemble (no reference counterpart: the reference runs one
 * mesh per process, examples/maxwell_2d/main.cpp) ----------------------------
 * n independent FP64 Maxwell TM instances (each its own discretization, size
 * and order) advanced together: one RK4 stage of all instances of one order
 * is ONE kernel launch over a work list of (instance, tile group) items; a
 * step (4 stages per order present, the orders as independent branches) is
 * replayed from a HIP graph. Instance i
 * uses operator ops[i], state vectors states[3i..3i+2] (Hx, Hy, Ez), work
 * vectors work[12i..12i+11] (4 groups of 3, as dgc_rk4_step's wk) and time
 * step dt[i]; all device vectors stay owned by the caller. */
typedef struct dgc_ensemble_s* dgc_ensemble_t;
int dgc_ensemble_create(dgc_ctx_t ctx, int n, const dgc_op_t* ops, double* const* states, double* const* work,
                        const double* dt, dgc_ensemble_t* out);
int dgc_ensemble_destroy(dgc_ensemble_t e);
/* Picks the tiling (64, 128 or 256 cells per workgroup) of each order present
 * by timing that order's largest instance (dgc_op_autotune; only work vectors
 * are written) and rebuilds the work list. */
int dgc_ensemble_autotune(dgc_ensemble_t e, int reps);
/* wt[n]: waves per 64-cell tile per instance; work-list length, launch grid, LDS bytes. */
int dgc_ensemble_info(dgc_ensemble_t e, int* wt, int* num_items, int* grid, size_t* lds_bytes);
/* Cache policy of the ensemble's stage kernels (DGC_CACHE_*; AUTO streams
 * when the working sets of all instances together exceed the last-level
 * cache). dgc_ensemble_cache_streaming: 1 if the streaming kernels run. */
int dgc_ensemble_set_cache_policy(dgc_ensemble_t e, int policy);
int dgc_ensemble_cache_streaming(dgc_ensemble_t e);
/* nsteps RK4 steps of every instance (graph replays); *ms: HIP-event time. */
int dgc_ensemble_steps(dgc_ensemble_t e, int nsteps, float* ms);
/* Baseline: the same steps instance after instance with the per-instance
 * kernels and step graphs (dgc_rk4_steps); *ms: HIP-event time. */
int dgc_ensemble_steps_sequential(dgc_ensemble_t e, int nsteps, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* DGC_HIP_H */
