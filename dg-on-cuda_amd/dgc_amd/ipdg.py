"""IPDG Poisson operator and solvers (ctypes over dgc_ipdg_* in include/dgc_hip.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .core import DeviceArray, Discretization, _d, _vp, check, lib, ref_operators


class IPDG:
    """laplace_2d on the GPU: Lap_h with homogeneous Dirichlet data.

    tau_mode 0 is the reference's tau = 1; tau_mode 1 is (N+1)(N+2)/h per face.
    """

    def __init__(self, disc: Discretization, tau_mode: int = 0, tau_scale: float = 1.0,
                 ops: dict | None = None):
        self.disc = disc
        self.ctx = disc.ctx
        ops = ops if ops is not None else ref_operators(disc.order, "strong")
        self.ops = ops
        h = _vp()
        check(lib().dgc_ipdg_create(disc.handle, _d(np.ascontiguousarray(ops["Dr"])),
                                    _d(np.ascontiguousarray(ops["Ds"])),
                                    _d(np.ascontiguousarray(ops["LIFT"])),
                                    _d(np.ascontiguousarray(ops["M"])), tau_mode, tau_scale, C.byref(h)))
        self._h = h
        self.num_colors = None

    @property
    def size(self):
        return self.disc.size

    def apply(self, u: DeviceArray):
        check(lib().dgc_ipdg_apply(self._h, _vp(u.ptr)))

    def apply_spd(self, u: DeviceArray, out: DeviceArray):
        check(lib().dgc_ipdg_apply_spd(self._h, _vp(u.ptr), _vp(out.ptr)))

    DIRICHLET, NEUMANN = 0, 1

    def set_boundary(self, kind: int):
        """Boundary kind of every boundary face (the BC object's is_dirichlet()):
        IPDG.DIRICHLET (default) or IPDG.NEUMANN. Drops the block-Jacobi
        inverses and the coarse space, which depend on it."""
        check(lib().dgc_ipdg_set_boundary(self._h, kind))
        self.num_colors = None

    def apply_inhomogeneous(self, inout: DeviceArray, gu: DeviceArray | None = None,
                            hq: DeviceArray | None = None):
        """inout += apply_inhomogeneous_as_rhs for boundary data per boundary node:
        gu = exterior_val(x, y, 0) on Dirichlet faces, hq = exterior_grad_n(x, y, 0)
        on Neumann faces (None: 0)."""
        check(lib().dgc_ipdg_apply_inhomogeneous(self._h, _vp(gu.ptr) if gu is not None else None,
                                                 _vp(hq.ptr) if hq is not None else None, _vp(inout.ptr)))

    def apply_data(self, u: DeviceArray, gd: DeviceArray | None = None, hd: DeviceArray | None = None):
        """In place u <- Lap_h u with boundary data (the whole BR1 operator):
        Dirichlet faces u+ = gd - u-, Neumann faces q+.n = hd - q-.n."""
        check(lib().dgc_ipdg_apply_data(self._h, _vp(gd.ptr) if gd is not None else None,
                                        _vp(hd.ptr) if hd is not None else None, _vp(u.ptr)))

    def set_graphs(self, enabled: bool):
        """Replay solver iterations from captured graphs (default) or launch the
        kernels one by one (for kernel tracing); the results are identical."""
        check(lib().dgc_ipdg_set_graphs(self._h, int(bool(enabled))))

    def set_shift(self, sigma: float):
        """SPD form -M Lap_h + sigma M for apply_spd, the block-Jacobi blocks and pcg."""
        check(lib().dgc_ipdg_set_shift(self._h, float(sigma)))

    def assemble(self):
        """The operator of apply() as (row_ptr, col_idx, values), CSR over the
        node-major DOF order (scipy.sparse.csr_matrix((values, col_idx, row_ptr)))."""
        n = self.disc.size
        nnz = C.c_longlong()
        rp = np.zeros(n + 1, np.int64)
        check(lib().dgc_ipdg_assemble(self._h, C.byref(nnz), rp.ctypes.data_as(C.POINTER(C.c_longlong)), None, None))
        ci = np.zeros(nnz.value, np.int32)
        va = np.zeros(nnz.value)
        check(lib().dgc_ipdg_assemble(self._h, C.byref(nnz), rp.ctypes.data_as(C.POINTER(C.c_longlong)),
                                      ci.ctypes.data_as(C.POINTER(C.c_int)), _d(va)))
        return rp, ci, va

    def mass(self, f: DeviceArray, out: DeviceArray):
        check(lib().dgc_ipdg_mass(self._h, _vp(f.ptr), _vp(out.ptr)))

    def block_jacobi(self) -> int:
        n = C.c_int()
        check(lib().dgc_ipdg_block_jacobi(self._h, C.byref(n)))
        self.num_colors = n.value
        return n.value

    def diag_blocks(self) -> np.ndarray:
        """Diagonal blocks of -M Lap_h, shape (K, Np, Np)."""
        K, Np = self.disc.K, self.disc.Np
        b = np.zeros(Np * Np * K)
        check(lib().dgc_ipdg_diag_blocks(self._h, _d(b)))
        return b.reshape(Np, Np, K).transpose(2, 0, 1).copy()

    def coarse_space(self, mesh, gx: int, gy: int | None = None) -> int:
        """Two-level coarse space: linear functions on gx x gy boxes of cells."""
        c = mesh.cells()
        n = C.c_int()
        check(lib().dgc_ipdg_coarse_space(self._h, _d(np.ascontiguousarray(c["cx"], np.float64)),
                                          _d(np.ascontiguousarray(c["cy"], np.float64)), gx,
                                          gy if gy is not None else gx, C.byref(n)))
        self.num_coarse = n.value
        return n.value

    def coarse_info(self) -> dict:
        nc, K, Np = self.num_coarse, self.disc.K, self.disc.Np
        A = np.zeros(nc * nc)
        Ainv = np.zeros(nc * nc, np.float32)
        agg = np.zeros(K, np.int32)
        coef = np.zeros(6 * K, np.float32)
        rs = np.zeros(2 * Np)
        check(lib().dgc_ipdg_coarse_info(self._h, _d(A), Ainv.ctypes.data_as(C.POINTER(C.c_float)),
                                         agg.ctypes.data_as(C.POINTER(C.c_int)),
                                         coef.ctypes.data_as(C.POINTER(C.c_float)), _d(rs)))
        return dict(A=A.reshape(nc, nc), Ainv=Ainv.reshape(nc, nc), cell_agg=agg, coef=coef.reshape(6, K),
                    r=rs[:Np], s=rs[Np:])

    def pcg(self, b: DeviceArray, x: DeviceArray, rtol=1e-10, maxit=20000, precond=1,
            history=False) -> dict:
        """precond 0: none, 1: block-Jacobi, 2: two-level (after coarse_space)."""
        it = C.c_int()
        rel = C.c_double()
        ms = C.c_float()
        hist = np.zeros(maxit + 1) if history else None
        rc = lib().dgc_ipdg_pcg(self._h, _vp(b.ptr), _vp(x.ptr), rtol, maxit, precond, C.byref(it),
                                C.byref(rel), _d(hist) if history else None, C.byref(ms))
        check(rc)
        out = dict(iters=it.value, relres=rel.value, ms=ms.value)
        if history:
            out["history"] = np.sqrt(hist[: it.value + 1])
        return out

    def pcg_mixed(self, b: DeviceArray, x: DeviceArray, rtol=1e-10, inner_rtol=1e-4, max_outer=30,
                  max_inner=20000) -> dict:
        o = C.c_int()
        i = C.c_int()
        rel = C.c_double()
        ms = C.c_float()
        check(lib().dgc_ipdg_pcg_mixed(self._h, _vp(b.ptr), _vp(x.ptr), rtol, inner_rtol, max_outer,
                                       max_inner, C.byref(o), C.byref(i), C.byref(rel), C.byref(ms)))
        return dict(outer=o.value, inner_iters=i.value, relres=rel.value, ms=ms.value)

    def __del__(self):
        if getattr(self, "_h", None) is not None and self.ctx._h is not None:
            lib().dgc_ipdg_destroy(self._h)
            self._h = None


def poisson_rhs(x, y, seed=3, nmodes=16):
    """Right-hand side of the Poisson benchmark problem on [-1,1]^2 and its exact solution.

    f = 2 pi^2 sin(pi x) sin(pi y) + sum over (m, n) in {1..4}^2 of
    a_mn sin(m pi (x+1)/2) sin(n pi (y+1)/2), a_mn ~ N(0,1) (m^2+n^2)^-1
    (amplitude decaying as k^-2 with k = |(m, n)|). The modes are Dirichlet
    eigenfunctions, so u = sin(pi x) sin(pi y) + sum a_mn / lambda_mn phi_mn.
    """
    rng = np.random.default_rng(seed)
    side = int(round(np.sqrt(nmodes)))
    f = 2 * np.pi**2 * np.sin(np.pi * x) * np.sin(np.pi * y)
    u = np.sin(np.pi * x) * np.sin(np.pi * y)
    for m in range(1, side + 1):
        for n in range(1, side + 1):
            k2 = m * m + n * n
            a = rng.standard_normal() / k2
            phi = np.sin(m * np.pi * (x + 1) / 2) * np.sin(n * np.pi * (y + 1) / 2)
            lam = (np.pi / 2) ** 2 * k2
            f = f + a * phi
            u = u + a / lam * phi
    return f, u
