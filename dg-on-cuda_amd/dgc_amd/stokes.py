"""Stokes dual-splitting step (ctypes over dgc_stokes_* in include/dgc_hip.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .core import DeviceArray, Discretization, _d, _i, _vp, check, lib

A = 2.883356
LAMBDA = 1 + A * A


def exact_velocity(x, y, t):
    """The example's exact solution (examples/stokes_2d/stokes_2d.h:297-298)."""
    ux = np.sin(x) * (A * np.sin(A * y) - np.cos(A) * np.sinh(y)) * np.exp(-LAMBDA * t)
    uy = np.cos(x) * (np.cos(A * y) + np.cos(A) * np.cosh(y)) * np.exp(-LAMBDA * t)
    return ux, uy


class Stokes:
    """stokes_2d::advance_timestep of the example problem on the GPU."""

    def __init__(self, disc: Discretization):
        self.disc = disc
        h = _vp()
        check(lib().dgc_stokes_create(disc.handle, C.byref(h)))
        self._h = h

    def coarse_space(self, mesh, gx: int, gy: int | None = None):
        """Two-level pressure solves over gx x gy boxes of cells."""
        c = mesh.cells()
        check(lib().dgc_stokes_coarse_space(self._h, _d(c["cx"]), _d(c["cy"]), gx, gy if gy else gx))

    def reset(self):
        """Forget the retained pressure (a new trajectory starts from p = 0)."""
        check(lib().dgc_stokes_reset(self._h))

    def advance(self, u_prev, t_prev, u, t, dt, out, rtol=1e-12) -> dict:
        """u_prev, u, out: pairs of DeviceArray (x, y components), node-major."""
        it = np.zeros(3, np.int32)
        check(lib().dgc_stokes_advance(self._h, _vp(u_prev[0].ptr), _vp(u_prev[1].ptr), t_prev, _vp(u[0].ptr),
                                       _vp(u[1].ptr), t, dt, rtol, _vp(out[0].ptr), _vp(out[1].ptr), _i(it)))
        return dict(pressure=int(it[0]), ux=int(it[1]), uy=int(it[2]))

    def __del__(self):
        if getattr(self, "_h", None) is not None and self.disc.ctx._h is not None:
            lib().dgc_stokes_destroy(self._h)
            self._h = None
