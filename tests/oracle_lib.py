"""ctypes access to the oracle (oracle/liboracle.so) -- test infrastructure.

The oracle is the CPU restatement of the reference's serial path; tests use
it as the checker of the HIP path, never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
_PATH_OMP = os.path.join(ROOT, "oracle", "liboracle_omp.so")  # threaded cell loops
_LIBS = {}
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def lib(threaded=False):
    path = _PATH_OMP if threaded else _PATH
    if path not in _LIBS:
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run oracle/build_ref.sh (or __graft_entry__.build())")
        L = C.CDLL(path)
        L.oracle_maxwell_rhs.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp,
                                         C.c_double, _dp]
        L.oracle_maxwell_rhs.restype = None
        L.oracle_advection_rhs.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp,
                                           _dp, C.c_double, C.c_double, C.c_int, _dp, C.c_double, _dp]
        L.oracle_advection_rhs.restype = None
        L.oracle_laplace_apply.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _ip, C.c_int,
                                           _dp]
        L.oracle_laplace_apply.restype = None
        L.oracle_rk4.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp, _dp,
                                 C.c_double, C.c_double, C.c_int, _dp, C.c_double, C.c_double, C.c_int,
                                 C.c_double]
        L.oracle_rk4.restype = C.c_int
        L.oracle_laplace_apply_bc.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _ip, C.c_int,
                                              C.c_int, _dp]
        L.oracle_laplace_apply_bc.restype = None
        L.oracle_laplace_apply_data.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _ip, C.c_int,
                                                C.c_int, _dp, _dp, _dp]
        L.oracle_laplace_apply_data.restype = None
        L.oracle_laplace_inhomogeneous.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _ip, C.c_int,
                                                   C.c_int, _dp, _dp, _dp]
        L.oracle_laplace_inhomogeneous.restype = None
        L.oracle_adv1d_rhs.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, _dp]
        L.oracle_adv1d_rhs.restype = None
        L.oracle_adv1d_rk4.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int]
        L.oracle_adv1d_rk4.restype = None
        L.oracle_grad.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp, C.c_double, _dp,
                                  _dp, _dp]
        L.oracle_grad.restype = None
        L.oracle_div.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp, _dp, C.c_double,
                                 _dp, _dp, _dp]
        L.oracle_div.restype = None
        _LIBS[path] = L
    return _LIBS[path]


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


class Problem:
    """Mesh + operators in the oracle's (reference CPU) conventions."""

    def __init__(self, order, Dr, Ds, L, cx, cy, nb_cell, nb_face, r=None, s=None, threaded=False):
        self._lib = lib(threaded)
        self.N = order
        self.Np = (order + 1) * (order + 2) // 2
        self.K = len(cx) // 3
        c = np.ascontiguousarray
        self.Dr, self.Ds, self.L = c(Dr, np.float64), c(Ds, np.float64), c(L, np.float64)
        self.cx, self.cy = c(cx, np.float64), c(cy, np.float64)
        self.nb_cell, self.nb_face = c(nb_cell, np.int32), c(nb_face, np.int32)
        self.r = c(r if r is not None else np.zeros(self.Np), np.float64)
        self.s = c(s if s is not None else np.zeros(self.Np), np.float64)

    @classmethod
    def from_mesh(cls, mesh, order, form="weak", ops=None, threaded=False):
        import dgc_amd

        ops = ops if ops is not None else dgc_amd.ref_operators(order, form)
        cells = mesh.cells()
        return cls(order, ops["Dr"], ops["Ds"], ops["LIFT"], cells["cx"], cells["cy"], cells["nb_cell"],
                   cells["nb_face"], ops["r"], ops["s"], threaded=threaded)

    def maxwell_rhs(self, u, t=0.0):
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros_like(u)
        self._lib.oracle_maxwell_rhs(self.K, self.N, _d(self.Dr), _d(self.Ds), _d(self.L), _d(self.cx),
                                 _d(self.cy), _i(self.nb_cell), _i(self.nb_face), _d(u), t, _d(out))
        return out

    def advection_rhs(self, u, t=0.0, ax=1.0, ay=1.0, bc=0):
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros_like(u)
        self._lib.oracle_advection_rhs(self.K, self.N, _d(self.Dr), _d(self.Ds), _d(self.L), _d(self.cx),
                                   _d(self.cy), _i(self.nb_cell), _i(self.nb_face), _d(self.r),
                                   _d(self.s), ax, ay, bc, _d(u), t, _d(out))
        return out

    def laplace(self, u, tau_mode=0):
        v = np.array(u, dtype=np.float64, copy=True)
        self._lib.oracle_laplace_apply(self.K, self.N, _d(self.Dr), _d(self.Ds), _d(self.L), _d(self.cx),
                                   _d(self.cy), _i(self.nb_cell), _i(self.nb_face), tau_mode, _d(v))
        return v

    def laplace_bc(self, u, neumann, tau_mode=0):
        """apply_homogeneous with one boundary kind (0 Dirichlet, 1 Neumann)."""
        v = np.array(u, dtype=np.float64, copy=True)
        self._lib.oracle_laplace_apply_bc(self.K, self.N, _d(self.Dr), _d(self.Ds), _d(self.L), _d(self.cx),
                                          _d(self.cy), _i(self.nb_cell), _i(self.nb_face), tau_mode, neumann, _d(v))
        return v

    def laplace_data(self, u, neumann, gd=None, hd=None, tau_mode=0):
        """The BR1 operator with boundary data (Dirichlet gd / Neumann hd per boundary node)."""
        v = np.array(u, dtype=np.float64, copy=True)
        gd = np.ascontiguousarray(gd, np.float64) if gd is not None else None
        hd = np.ascontiguousarray(hd, np.float64) if hd is not None else None
        self._lib.oracle_laplace_apply_data(self.K, self.N, _d(self.Dr), _d(self.Ds), _d(self.L), _d(self.cx),
                                            _d(self.cy), _i(self.nb_cell), _i(self.nb_face), tau_mode, neumann,
                                            _d(gd) if gd is not None else None,
                                            _d(hd) if hd is not None else None, _d(v))
        return v

    def laplace_inhomogeneous(self, inout, neumann, gu=None, hq=None, tau_mode=0):
        """inout + apply_inhomogeneous_as_rhs for boundary data gu (Dirichlet) or hq (Neumann)."""
        v = np.array(inout, dtype=np.float64, copy=True)
        gu = np.ascontiguousarray(gu, np.float64) if gu is not None else None
        hq = np.ascontiguousarray(hq, np.float64) if hq is not None else None
        self._lib.oracle_laplace_inhomogeneous(self.K, self.N, _d(self.Dr), _d(self.Ds), _d(self.L), _d(self.cx),
                                               _d(self.cy), _i(self.nb_cell), _i(self.nb_face), tau_mode, neumann,
                                               _d(gu) if gu is not None else None,
                                               _d(hq) if hq is not None else None, _d(v))
        return v

    def grad(self, u, g=None, beta=1.0):
        """Weak gradient (grad_2d) with exterior value g[b] + beta * interior."""
        u = np.ascontiguousarray(u, np.float64)
        ox, oy = np.zeros_like(u), np.zeros_like(u)
        gp = _d(np.ascontiguousarray(g, np.float64)) if g is not None else None
        self._lib.oracle_grad(self.K, self.N, _d(self.Dr), _d(self.Ds), _d(self.L), _d(self.cx), _d(self.cy),
                              _i(self.nb_cell), _i(self.nb_face), gp, beta, _d(u), _d(ox), _d(oy))
        return ox, oy

    def div(self, u0, u1, g0=None, g1=None, beta=1.0):
        """Weak divergence (div_2d) with exterior values g_i[b] + beta * interior."""
        u0 = np.ascontiguousarray(u0, np.float64)
        u1 = np.ascontiguousarray(u1, np.float64)
        out = np.zeros_like(u0)
        g0 = np.ascontiguousarray(g0, np.float64) if g0 is not None else None
        g1 = np.ascontiguousarray(g1, np.float64) if g1 is not None else None
        self._lib.oracle_div(self.K, self.N, _d(self.Dr), _d(self.Ds), _d(self.L), _d(self.cx), _d(self.cy),
                             _i(self.nb_cell), _i(self.nb_face), _d(g0) if g0 is not None else None,
                             _d(g1) if g1 is not None else None, beta, _d(u0), _d(u1), _d(out))
        return out

    def rk4(self, phys, u, t0, dt, nsteps=0, t_end=-1.0, ax=1.0, ay=1.0, bc=0):
        v = np.array(u, dtype=np.float64, copy=True)
        steps = self._lib.oracle_rk4(phys, self.K, self.N, _d(self.Dr), _d(self.Ds), _d(self.L),
                                 _d(self.cx), _d(self.cy), _i(self.nb_cell), _i(self.nb_face),
                                 _d(self.r), _d(self.s), ax, ay, bc, _d(v), t0, dt, nsteps, t_end)
        return v, steps


def adv1d_rhs(K, N, D, L, u, t):
    """advection_1d::operator() restated (cell-major)."""
    D = np.ascontiguousarray(D, np.float64)
    L = np.ascontiguousarray(L, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    out = np.zeros_like(u)
    lib().oracle_adv1d_rhs(K, N, _d(D), _d(L), _d(u), t, _d(out))
    return out


def adv1d_rk4(K, N, D, L, u, t0, dt, nsteps):
    D = np.ascontiguousarray(D, np.float64)
    L = np.ascontiguousarray(L, np.float64)
    v = np.array(u, dtype=np.float64, copy=True)
    lib().oracle_adv1d_rk4(K, N, _d(D), _d(L), _d(v), t0, dt, nsteps)
    return v
