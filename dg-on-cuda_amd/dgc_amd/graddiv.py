"""Gradient and divergence with central fluxes (ctypes over dgc_graddiv_*,
dgc_grad2d_apply and dgc_div2d_apply in include/dgc_hip.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .core import DeviceArray, Discretization, _d, _vp, check, lib, ref_operators


def _ptr(a: DeviceArray | None):
    return _vp(a.ptr) if a is not None else None


class GradDiv:
    """grad_2d and div_2d (src/grad_2d.h, src/div_2d.h) on the GPU.

    Boundary data: the exterior value at boundary node b is g[b] + beta * the
    interior value, boundary nodes numbered as Discretization.arrays()["bx"]
    (g = None: 0). beta = 1 with g = None is the "exterior = interior" BC;
    g = 2 u_D, beta = -1 imposes the Dirichlet value u_D through the central
    flux. Outputs must not alias inputs.
    """

    def __init__(self, disc: Discretization, ops: dict | None = None):
        self.disc = disc
        ops = ops if ops is not None else ref_operators(disc.order, "weak")
        h = _vp()
        check(lib().dgc_graddiv_create(disc.handle, _d(np.ascontiguousarray(ops["Dr"])),
                                       _d(np.ascontiguousarray(ops["Ds"])),
                                       _d(np.ascontiguousarray(ops["LIFT"])), C.byref(h)))
        self._h = h

    def set_tiling(self, waves_per_tile: int):
        """1, 2 or 4 waves per 64-cell tile (0: the order's default)."""
        check(lib().dgc_graddiv_set_tiling(self._h, waves_per_tile))

    def grad(self, u: DeviceArray, out_x: DeviceArray, out_y: DeviceArray, g: DeviceArray | None = None,
             beta: float = 1.0):
        check(lib().dgc_grad2d_apply(self._h, _vp(u.ptr), _ptr(g), beta, _vp(out_x.ptr), _vp(out_y.ptr)))

    def div(self, u0: DeviceArray, u1: DeviceArray, out: DeviceArray, g0: DeviceArray | None = None,
            g1: DeviceArray | None = None, beta: float = 1.0):
        check(lib().dgc_div2d_apply(self._h, _vp(u0.ptr), _vp(u1.ptr), _ptr(g0), _ptr(g1), beta, _vp(out.ptr)))

    def __del__(self):
        if getattr(self, "_h", None) is not None and self.disc.ctx._h is not None:
            lib().dgc_graddiv_destroy(self._h)
            self._h = None
