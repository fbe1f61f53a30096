"""Batched ensemble of Maxwell TM instances (ctypes over dgc_ensemble_*)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .core import DeviceArray, _vp, check, lib

_dp = C.POINTER(C.c_double)


def _declare():
    L = lib()
    if getattr(L, "_ens_declared", False):
        return L

    def d(name, args, res=C.c_int):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res

    d("dgc_ensemble_create", [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), _dp, C.POINTER(_vp)])
    d("dgc_ensemble_destroy", [_vp])
    d("dgc_ensemble_autotune", [_vp, C.c_int])
    d("dgc_ensemble_info", [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                            C.POINTER(C.c_size_t)])
    d("dgc_ensemble_set_cache_policy", [_vp, C.c_int])
    d("dgc_ensemble_cache_streaming", [_vp])
    d("dgc_ensemble_steps", [_vp, C.c_int, C.POINTER(C.c_float)])
    d("dgc_ensemble_steps_sequential", [_vp, C.c_int, C.POINTER(C.c_float)])
    L._ens_declared = True
    return L


class Ensemble:
    """Independent Maxwell TM instances advanced together: one launch per
    order and RK4 stage for all instances of that order, the step replayed
    from a HIP graph.

    ops[i]: MaxwellTM operator, states[i]: its 3 field DeviceArrays,
    works[i]: its work vectors (op.new_work()), dts[i]: its time step."""

    def __init__(self, ctx, ops, states, works, dts):
        L = _declare()
        self.ctx = ctx
        self.ops, self.states, self.works = list(ops), list(states), list(works)
        n = len(self.ops)
        sp = (_vp * (3 * n))(*[_vp(a.ptr) for s in self.states for a in s])
        wp = (_vp * (12 * n))(*[_vp(a.ptr) for w in self.works for a in w[:12]])
        op = (_vp * n)(*[o._h for o in self.ops])
        self.dts = np.ascontiguousarray(dts, np.float64)
        h = _vp()
        check(L.dgc_ensemble_create(ctx.handle, n, op, sp, wp, self.dts.ctypes.data_as(_dp), C.byref(h)))
        self._h = h
        self.n = n

    def autotune(self, reps: int = 20):
        check(lib().dgc_ensemble_autotune(self._h, reps))

    def info(self) -> dict:
        wt = (C.c_int * self.n)()
        ni, g = C.c_int(), C.c_int()
        lds = C.c_size_t()
        check(lib().dgc_ensemble_info(self._h, wt, C.byref(ni), C.byref(g), C.byref(lds)))
        return dict(waves_per_tile=list(wt), num_items=ni.value, grid=g.value, lds_bytes=lds.value)

    @property
    def streaming(self) -> bool:
        """True if the stage kernels run with nontemporal accumulator streams."""
        return lib().dgc_ensemble_cache_streaming(self._h) == 1

    def set_cache_policy(self, policy: str):
        """'auto' (stream when all instances' working sets together exceed the
        256 MB last-level cache), 'temporal' or 'streaming'."""
        check(lib().dgc_ensemble_set_cache_policy(self._h, {"auto": 0, "temporal": 1, "streaming": 2}[policy]))

    def steps(self, nsteps: int) -> float:
        ms = C.c_float()
        check(lib().dgc_ensemble_steps(self._h, nsteps, C.byref(ms)))
        return ms.value

    def steps_sequential(self, nsteps: int) -> float:
        ms = C.c_float()
        check(lib().dgc_ensemble_steps_sequential(self._h, nsteps, C.byref(ms)))
        return ms.value

    def __del__(self):
        if getattr(self, "_h", None) is not None and self.ctx._h is not None:
            lib().dgc_ensemble_destroy(self._h)
            self._h = None
