"""Learned preconditioner and flexible CG (ctypes over dgc_mpnn_* / dgc_ipdg_fcg)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .core import DeviceArray, _vp, check, lib

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_NOT_CONVERGED = 5  # DGC_ERR_NOT_CONVERGED


def _declare():
    L = lib()
    if getattr(L, "_mpnn_declared", False):
        return L

    def d(name, args, res=C.c_int):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res

    d("dgc_mpnn_create", [_vp, C.c_uint64, C.c_int, C.POINTER(_vp)])
    d("dgc_mpnn_destroy", [_vp])
    d("dgc_mpnn_param_count", [_vp])
    d("dgc_mpnn_input_width", [_vp])
    d("dgc_mpnn_get_params", [_vp, _fp])
    d("dgc_mpnn_set_params", [_vp, _fp])
    d("dgc_mpnn_eval", [_vp, _fp, _fp])
    d("dgc_mpnn_loss_grad", [_vp, _fp, _fp, _dp, _fp])
    d("dgc_mpnn_build_dataset", [_vp, _dp, _dp, C.c_int, C.c_uint64, C.c_int])
    d("dgc_mpnn_build_dataset_solver", [_vp, _dp, _dp, C.c_int, C.c_uint64, C.c_int, C.c_int,
                                         C.POINTER(C.c_int)])
    d("dgc_mpnn_train", [_vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint64, _dp])
    d("dgc_mpnn_eval_loss", [_vp, C.c_int, C.c_int, _dp])
    d("dgc_mpnn_target_scale", [_vp, _dp])
    d("dgc_mpnn_set_target_scale", [_vp, C.c_double])
    d("dgc_mpnn_energy_loss_grad", [_vp, _vp, C.c_double, _dp, _fp, _vp])
    d("dgc_ipdg_fcg", [_vp, _vp, _vp, _vp, C.c_double, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_int), _dp,
                       C.POINTER(C.c_int), _dp, C.POINTER(C.c_float)])
    L._mpnn_declared = True
    return L


class LearnedPreconditioner:
    """3-layer message-passing network correcting a base preconditioner step:
    base 1 = block-Jacobi, base 2 = two-level (needs ipdg.coarse_space first)."""

    def __init__(self, ipdg, seed: int = 0, base: int = 1):
        L = _declare()
        self.ipdg = ipdg
        self.ctx = ipdg.ctx
        self.base = base
        h = _vp()
        check(L.dgc_mpnn_create(ipdg._h, seed, base, C.byref(h)))
        self._h = h
        self.P = L.dgc_mpnn_param_count(h)
        self.D0 = L.dgc_mpnn_input_width(h)
        self.K = ipdg.disc.K
        self.Np = ipdg.disc.Np

    def get_params(self) -> np.ndarray:
        w = np.zeros(self.P, np.float32)
        check(lib().dgc_mpnn_get_params(self._h, w.ctypes.data_as(_fp)))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.P
        check(lib().dgc_mpnn_set_params(self._h, w.ctypes.data_as(_fp)))

    def eval(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32).reshape(self.K, self.D0)
        y = np.zeros((self.K, self.Np), np.float32)
        check(lib().dgc_mpnn_eval(self._h, X.ctypes.data_as(_fp), y.ctypes.data_as(_fp)))
        return y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, np.float32).reshape(self.K, self.D0)
        T = np.ascontiguousarray(T, np.float32).reshape(self.K, self.Np)
        loss = C.c_double()
        g = np.zeros(self.P, np.float32)
        check(lib().dgc_mpnn_loss_grad(self._h, X.ctypes.data_as(_fp), T.ctypes.data_as(_fp), C.byref(loss),
                                       g.ctypes.data_as(_fp)))
        return loss.value, g

    @property
    def target_scale(self) -> float:
        v = C.c_double()
        check(lib().dgc_mpnn_target_scale(self._h, C.byref(v)))
        return v.value

    @target_scale.setter
    def target_scale(self, ct: float):
        check(lib().dgc_mpnn_set_target_scale(self._h, float(ct)))

    def energy_loss_grad(self, r: DeviceArray, er: float, z: DeviceArray | None = None):
        """Energy-norm loss |z - e|_S^2 / |e|_S^2 of the residual r = S e (er = e.r)
        and its parameter gradient; z (optional) receives the preconditioned vector."""
        loss = C.c_double()
        g = np.zeros(self.P, np.float32)
        check(lib().dgc_mpnn_energy_loss_grad(self._h, _vp(r.ptr), float(er), C.byref(loss),
                                              g.ctypes.data_as(_fp), _vp(z.ptr) if z is not None else None))
        return loss.value, g

    def build_dataset(self, mesh, nsamples: int, seed0: int = 100, modes: int = 16):
        x, y = mesh.dof_coords(self.ipdg.disc.order, node_major=True)
        check(lib().dgc_mpnn_build_dataset(self._h, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp), nsamples,
                                           seed0, modes))
        self.nsamples = nsamples

    def build_dataset_solver(self, mesh, nrhs: int, seed0: int = 100, modes: int = 16,
                             snapshots: int = 16) -> int:
        """Pairs (r = S e, e = x* - x_j) from the iterates of base-preconditioned
        PCG solves of nrhs random right-hand sides; returns the pair count."""
        x, y = mesh.dof_coords(self.ipdg.disc.order, node_major=True)
        n = C.c_int()
        check(lib().dgc_mpnn_build_dataset_solver(self._h, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp), nrhs,
                                                  seed0, modes, snapshots, C.byref(n)))
        self.nsamples = n.value
        return n.value

    def train(self, first: int, count: int, epochs: int, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        hist = np.zeros(epochs)
        check(lib().dgc_mpnn_train(self._h, first, count, epochs, lr, seed, hist.ctypes.data_as(_dp)))
        return hist

    def eval_loss(self, first: int, count: int) -> float:
        v = C.c_double()
        check(lib().dgc_mpnn_eval_loss(self._h, first, count, C.byref(v)))
        return v.value

    def fcg(self, b: DeviceArray, x: DeviceArray, rtol=1e-10, maxit=20000, stall_window=1000,
            stall_factor=0.99, history=False, raise_on_cap=True) -> dict:
        """Flexible CG; with raise_on_cap=False an iteration cap is reported as
        converged=False instead of raising."""
        it = C.c_int()
        rel = C.c_double()
        fb = C.c_int()
        ms = C.c_float()
        hist = np.zeros(maxit + 1) if history else None
        rc = lib().dgc_ipdg_fcg(self.ipdg._h, self._h, _vp(b.ptr), _vp(x.ptr), rtol, maxit, stall_window,
                                stall_factor, C.byref(it), C.byref(rel), C.byref(fb),
                                hist.ctypes.data_as(_dp) if history else None, C.byref(ms))
        if raise_on_cap or rc != _NOT_CONVERGED:
            check(rc)
        out = dict(iters=it.value, relres=rel.value, fell_back=fb.value, ms=ms.value, converged=rc == 0)
        if history:
            out["history"] = np.sqrt(hist[: it.value + 1])
        return out

    def __del__(self):
        if getattr(self, "_h", None) is not None and self.ctx._h is not None:
            lib().dgc_mpnn_destroy(self._h)
            self._h = None
