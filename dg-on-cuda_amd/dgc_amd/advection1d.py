"""1D advection (examples/advection_1d) and the examples' plot files
(ctypes over dgc_adv1d_* and dgc_write_plot_file in include/dgc_hip.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .core import Context, DeviceArray, _d, _vp, check, lib

WAVE_SPEED = 2 * np.pi
DOMAIN = 2 * np.pi


class Advection1D:
    """advection_1d on the GPU: u_t + 2 pi u_x = 0 on [0, 2 pi], inflow -sin(2 pi t)."""

    def __init__(self, ctx: Context, num_cells: int, order: int, D=None, L=None):
        self.ctx, self.K, self.order = ctx, num_cells, order
        self.Np = order + 1
        h = _vp()
        Dp = _d(np.ascontiguousarray(D, np.float64)) if D is not None else None
        Lp = _d(np.ascontiguousarray(L, np.float64)) if L is not None else None
        check(lib().dgc_adv1d_create(ctx.handle, num_cells, order, Dp, Lp, C.byref(h)))
        self._h = h

    @staticmethod
    def operators(num_cells, order):
        """(D, L, x): strong-form D = Dr 2/h, L = V V^T 2/h, node positions (node-major)."""
        Np = order + 1
        D, L, x = np.zeros(Np * Np), np.zeros(Np * Np), np.zeros(num_cells * Np)
        check(lib().dgc_adv1d_operators(num_cells, order, _d(D), _d(L), _d(x)))
        return D.reshape(Np, Np), L.reshape(Np, Np), x

    @property
    def size(self):
        return self.K * self.Np

    def min_elem_size(self):
        return DOMAIN / self.K

    def rhs(self, u: DeviceArray, t: float, out: DeviceArray):
        check(lib().dgc_adv1d_rhs(self._h, _vp(u.ptr), t, _vp(out.ptr)))

    def rk4_steps(self, u: DeviceArray, t0: float, dt: float, nsteps: int):
        check(lib().dgc_adv1d_rk4_steps(self._h, _vp(u.ptr), t0, dt, nsteps))

    def __del__(self):
        if getattr(self, "_h", None) is not None and self.ctx._h is not None:
            lib().dgc_adv1d_destroy(self._h)
            self._h = None


def write_plot_file(path, x, u, ref, y=None):
    """The examples' plot data file (2D with y, else the 1D layout)."""
    x = np.ascontiguousarray(x, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    ref = np.ascontiguousarray(ref, np.float64)
    yp = _d(np.ascontiguousarray(y, np.float64)) if y is not None else None
    check(lib().dgc_write_plot_file(str(path).encode(), len(x), _d(x), yp, _d(u), _d(ref)))
