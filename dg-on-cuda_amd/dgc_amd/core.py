"""ctypes layer over libdgc_amd.so (include/dgc_hip.h)."""
from __future__ import annotations

import ctypes as C
import os
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def lib_path() -> str:
    return os.path.join(os.path.dirname(_HERE), "lib", "libdgc_amd.so")


class DGCError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[dgc status {code}] {msg}")
        self.code = code


_LIB = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


def _declare(L):
    def d(name, args, res=C.c_int):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res

    d("dgc_last_error", [], C.c_char_p)
    d("dgc_max_order", [])
    d("dgc_ctx_create", [C.c_int, C.POINTER(_vp)])
    d("dgc_ctx_destroy", [_vp])
    d("dgc_ctx_sync", [_vp])
    d("dgc_malloc", [_vp, C.c_size_t, C.POINTER(_vp)])
    d("dgc_free", [_vp, _vp])
    d("dgc_memcpy_h2d", [_vp, _vp, _vp, C.c_size_t])
    d("dgc_memcpy_d2h", [_vp, _vp, _vp, C.c_size_t])
    d("dgc_memcpy_d2d", [_vp, _vp, _vp, C.c_size_t])
    d("dgc_memset_zero", [_vp, _vp, C.c_size_t])
    d("dgc_ref_operators", [C.c_int, C.c_int, _dp, _dp, _ip, _dp, _dp, _dp, _dp])
    d("dgc_mesh_structured", [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                              C.c_double, C.c_uint64, C.c_int, C.POINTER(_vp)])
    d("dgc_mesh_from_arrays", [C.c_int, _dp, _dp, C.c_int, _ip, C.POINTER(_vp)])
    d("dgc_mesh_read_gmsh", [C.c_char_p, C.POINTER(_vp)])
    d("dgc_mesh_write_gmsh", [_vp, C.c_char_p])
    d("dgc_mesh_destroy", [_vp])
    d("dgc_mesh_info", [_vp, _ip, _ip, _dp])
    d("dgc_mesh_cells", [_vp, _dp, _dp, _ip, _ip, _ip])
    d("dgc_mesh_geometry", [_vp, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _dp, _dp])
    d("dgc_mesh_dof_coords", [_vp, C.c_int, C.c_int, _dp, _dp])
    d("dgc_disc2d_create", [_vp, C.c_int, C.c_int, _ip, _ip, _ip, _dp, _dp, _dp, _ip, _ip, C.c_int,
                            _dp, _dp, _dp, _dp, C.POINTER(_vp)])
    d("dgc_disc2d_destroy", [_vp])
    d("dgc_disc2d_create_from_cells", [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp, _ip, C.POINTER(_vp)])
    d("dgc_disc2d_arrays", [_vp, _dp, _ip, _dp, _dp, _ip, _dp, _dp])
    d("dgc_graddiv_create", [_vp, _dp, _dp, _dp, C.POINTER(_vp)])
    d("dgc_graddiv_destroy", [_vp])
    d("dgc_graddiv_set_tiling", [_vp, C.c_int])
    d("dgc_grad2d_apply", [_vp, _vp, _vp, C.c_double, _vp, _vp])
    d("dgc_div2d_apply", [_vp, _vp, _vp, _vp, _vp, C.c_double, _vp])
    d("dgc_stokes_create", [_vp, C.POINTER(_vp)])
    d("dgc_adv1d_operators", [C.c_int, C.c_int, _dp, _dp, _dp])
    d("dgc_adv1d_create", [_vp, C.c_int, C.c_int, _dp, _dp, C.POINTER(_vp)])
    d("dgc_adv1d_destroy", [_vp])
    d("dgc_adv1d_rhs", [_vp, _vp, C.c_double, _vp])
    d("dgc_adv1d_rk4_steps", [_vp, _vp, C.c_double, C.c_double, C.c_int])
    d("dgc_write_plot_file", [C.c_char_p, C.c_size_t, _dp, _dp, _dp, _dp])
    d("dgc_stokes_destroy", [_vp])
    d("dgc_stokes_reset", [_vp])
    d("dgc_stokes_coarse_space", [_vp, _dp, _dp, C.c_int, C.c_int])
    d("dgc_stokes_advance", [_vp, _vp, _vp, C.c_double, _vp, _vp, C.c_double, C.c_double, C.c_double, _vp, _vp,
                             _ip])
    d("dgc_maxwell_tm_create", [_vp, _dp, _dp, _dp, C.POINTER(_vp)])
    d("dgc_advection_create", [_vp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int, C.POINTER(_vp)])
    d("dgc_op_destroy", [_vp])
    d("dgc_op_num_fields", [_vp])
    d("dgc_op_rhs", [_vp, C.POINTER(_vp), C.c_size_t, C.c_double, C.POINTER(_vp)])
    d("dgc_axpy", [_vp, C.c_double, _vp, C.c_size_t, _vp, _vp])
    d("dgc_rk4_step", [_vp, C.POINTER(_vp), C.c_size_t, C.c_double, C.c_double, C.POINTER(_vp)])
    d("dgc_rk4_run", [_vp, C.POINTER(_vp), C.c_size_t, C.c_double, C.c_double, C.c_double,
                      C.POINTER(_vp), _dp, _ip])
    d("dgc_rk4_steps", [_vp, C.POINTER(_vp), C.c_size_t, C.c_double, C.c_double, C.c_int,
                        C.POINTER(_vp)])
    d("dgc_rk4_steps_timed", [_vp, C.POINTER(_vp), C.c_size_t, C.c_double, C.c_double, C.c_int,
                              C.POINTER(_vp), C.POINTER(C.c_float)])
    d("dgc_rk4_profile_stages", [_vp, C.POINTER(_vp), C.c_size_t, C.c_double, C.c_double, C.c_int,
                                 C.POINTER(_vp), C.POINTER(C.c_float)])
    d("dgc_ipdg_create", [_vp, _dp, _dp, _dp, _dp, C.c_int, C.c_double, C.POINTER(_vp)])
    d("dgc_ipdg_set_boundary", [_vp, C.c_int])
    d("dgc_ipdg_apply_inhomogeneous", [_vp, _vp, _vp, _vp])
    d("dgc_ipdg_apply_data", [_vp, _vp, _vp, _vp])
    d("dgc_ipdg_set_shift", [_vp, C.c_double])
    d("dgc_ipdg_set_graphs", [_vp, C.c_int])
    d("dgc_ipdg_assemble", [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), _ip, _dp])
    d("dgc_ipdg_destroy", [_vp])
    d("dgc_ipdg_apply", [_vp, _vp])
    d("dgc_ipdg_apply_spd", [_vp, _vp, _vp])
    d("dgc_ipdg_mass", [_vp, _vp, _vp])
    d("dgc_ipdg_block_jacobi", [_vp, _ip])
    d("dgc_ipdg_diag_blocks", [_vp, _dp])
    d("dgc_ipdg_coarse_space", [_vp, _dp, _dp, C.c_int, C.c_int, _ip])
    d("dgc_ipdg_coarse_info", [_vp, _dp, C.POINTER(C.c_float), _ip, C.POINTER(C.c_float), _dp])
    d("dgc_ipdg_pcg", [_vp, _vp, _vp, C.c_double, C.c_int, C.c_int, _ip, _dp, _dp,
                       C.POINTER(C.c_float)])
    d("dgc_ipdg_pcg_mixed", [_vp, _vp, _vp, C.c_double, C.c_double, C.c_int, C.c_int, _ip, _ip, _dp,
                             C.POINTER(C.c_float)])
    d("dgc_op_set_waves_per_tile", [_vp, C.c_int])
    d("dgc_op_waves_per_tile", [_vp])
    d("dgc_op_set_precision", [_vp, C.c_int])
    d("dgc_op_precision", [_vp])
    d("dgc_op_set_cache_policy", [_vp, C.c_int])
    d("dgc_op_cache_streaming", [_vp])
    d("dgc_op_autotune", [_vp, C.POINTER(_vp), C.c_size_t, C.POINTER(_vp), C.c_int, C.POINTER(C.c_float)])
    d("dgc_op_stage_time", [_vp, C.POINTER(_vp), C.c_size_t, C.POINTER(_vp), C.c_int, C.c_int,
                            C.POINTER(C.c_float)])
    for name, args in _EXTRA_DECLS:
        d(name, args)


# Declarations added by other modules (name, argtypes), applied on load.
_EXTRA_DECLS: list = []


def lib():
    """Load libdgc_amd.so once; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        p = lib_path()
        if not os.path.exists(p):
            raise DGCError(-1, f"HIP library not built: {p} (run __graft_entry__.build())")
        L = C.CDLL(p)
        _declare(L)
        _LIB = L
    return _LIB


def check(rc: int):
    if rc != 0:
        raise DGCError(rc, lib().dgc_last_error().decode())


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def _ptrs(values: Sequence[int]):
    arr = (_vp * len(values))(*values)
    return arr


# ---------------------------------------------------------------------------
def ref_operators(order: int, form: str = "weak") -> dict:
    """Nodes, face node lists, Dr, Ds, LIFT and mass of the reference triangle."""
    Np = (order + 1) * (order + 2) // 2
    Nfp = order + 1
    r = np.zeros(Np)
    s = np.zeros(Np)
    fm = np.zeros(3 * Nfp, dtype=np.int32)
    Dr = np.zeros((Np, Np))
    Ds = np.zeros((Np, Np))
    L = np.zeros((Np, 3 * Nfp))
    M = np.zeros((Np, Np))
    check(lib().dgc_ref_operators(order, 0 if form == "weak" else 1, _d(r), _d(s), _i(fm), _d(Dr),
                                  _d(Ds), _d(L), _d(M)))
    return dict(order=order, Np=Np, Nfp=Nfp, r=r, s=s, fmask=fm.reshape(3, Nfp), Dr=Dr, Ds=Ds,
                LIFT=L, M=M)


class Mesh:
    """Triangle mesh with half-edge topology (host set-up, C++)."""

    def __init__(self, handle):
        self._h = _vp(handle)
        K = C.c_int()
        Nv = C.c_int()
        hmin = C.c_double()
        check(lib().dgc_mesh_info(self._h, C.byref(K), C.byref(Nv), C.byref(hmin)))
        self.K = K.value
        self.Nv = Nv.value
        self.min_edge = hmin.value

    @classmethod
    def structured(cls, nx, ny, x0=-1.0, x1=1.0, y0=-1.0, y1=1.0, jitter=0.0, seed=0, periodic=False):
        h = _vp()
        check(lib().dgc_mesh_structured(nx, ny, x0, x1, y0, y1, jitter, seed, int(periodic), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_arrays(cls, vx, vy, etov):
        vx = np.ascontiguousarray(vx, dtype=np.float64)
        vy = np.ascontiguousarray(vy, dtype=np.float64)
        etov = np.ascontiguousarray(etov, dtype=np.int32).reshape(-1)
        h = _vp()
        check(lib().dgc_mesh_from_arrays(len(vx), _d(vx), _d(vy), len(etov) // 3, _i(etov), C.byref(h)))
        return cls(h.value)

    @classmethod
    def read_gmsh(cls, path):
        h = _vp()
        check(lib().dgc_mesh_read_gmsh(str(path).encode(), C.byref(h)))
        return cls(h.value)

    def write_gmsh(self, path):
        """gmsh MSH 4.1 ASCII (one node block, one triangle block)."""
        check(lib().dgc_mesh_write_gmsh(self._h, str(path).encode()))

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.dgc_mesh_destroy(self._h)
            self._h = None

    def cells(self) -> dict:
        n = 3 * self.K
        cx = np.zeros(n)
        cy = np.zeros(n)
        ev = np.zeros(n, np.int32)
        nc = np.zeros(n, np.int32)
        nf = np.zeros(n, np.int32)
        check(lib().dgc_mesh_cells(self._h, _d(cx), _d(cy), _i(ev), _i(nc), _i(nf)))
        return dict(cx=cx, cy=cy, etov=ev, nb_cell=nc, nb_face=nf)

    def vertices(self):
        c = self.cells()
        vx = np.zeros(self.Nv)
        vy = np.zeros(self.Nv)
        vx[c["etov"]] = c["cx"]
        vy[c["etov"]] = c["cy"]
        return vx, vy, c["etov"].reshape(-1, 3)

    def geometry(self, order: int) -> dict:
        K = self.K
        nb = C.c_int()
        g = dict(inv_jac=np.zeros(4 * K), J=np.zeros(K), face_J=np.zeros(3 * K), nx=np.zeros(3 * K),
                 ny=np.zeros(3 * K), iface_cell=np.zeros(3 * K, np.int32),
                 iface_face=np.zeros(3 * K, np.int32))
        check(lib().dgc_mesh_geometry(self._h, order, _d(g["inv_jac"]), _d(g["J"]), _d(g["face_J"]),
                                      _d(g["nx"]), _d(g["ny"]), _i(g["iface_cell"]),
                                      _i(g["iface_face"]), C.byref(nb), None, None))
        g["bx"] = np.zeros(max(nb.value, 1))
        g["by"] = np.zeros(max(nb.value, 1))
        check(lib().dgc_mesh_geometry(self._h, order, None, None, None, None, None, None, None,
                                      C.byref(nb), _d(g["bx"]), _d(g["by"])))
        g["num_boundary_nodes"] = nb.value
        g["bx"] = g["bx"][: nb.value]
        g["by"] = g["by"][: nb.value]
        return g

    def dof_coords(self, order: int, node_major: bool = True):
        Np = (order + 1) * (order + 2) // 2
        x = np.zeros(self.K * Np)
        y = np.zeros(self.K * Np)
        check(lib().dgc_mesh_dof_coords(self._h, order, int(node_major), _d(x), _d(y)))
        return x, y


def node_to_cell_major(u: np.ndarray, K: int, Np: int) -> np.ndarray:
    return np.ascontiguousarray(u.reshape(Np, K).T).reshape(-1)


def cell_to_node_major(u: np.ndarray, K: int, Np: int) -> np.ndarray:
    return np.ascontiguousarray(u.reshape(K, Np).T).reshape(-1)


# ---------------------------------------------------------------------------
class Context:
    def __init__(self, device: int = 0):
        h = _vp()
        check(lib().dgc_ctx_create(device, C.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def sync(self):
        check(lib().dgc_ctx_sync(self._h))

    def close(self):
        if self._h is not None:
            lib().dgc_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.dgc_ctx_destroy(self._h)
            self._h = None


class DeviceArray:
    """A device buffer of float64 (or int32) owned through the C ABI."""

    def __init__(self, ctx: Context, n: int, dtype=np.float64):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = _vp()
        check(lib().dgc_malloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p.value

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    @classmethod
    def from_host(cls, ctx: Context, a: np.ndarray):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.n:
            raise ValueError("size mismatch")
        check(lib().dgc_memcpy_h2d(self.ctx.handle, _vp(self.ptr), a.ctypes.data_as(_vp), self.nbytes))

    def download(self) -> np.ndarray:
        a = np.empty(self.n, dtype=self.dtype)
        check(lib().dgc_memcpy_d2h(self.ctx.handle, a.ctypes.data_as(_vp), _vp(self.ptr), self.nbytes))
        return a

    def zero(self):
        check(lib().dgc_memset_zero(self.ctx.handle, _vp(self.ptr), self.nbytes))

    def free(self):
        if self.ptr:
            lib().dgc_free(self.ctx.handle, _vp(self.ptr))
            self.ptr = 0

    def __del__(self):
        if getattr(self, "ptr", 0) and _LIB is not None and self.ctx._h is not None:
            _LIB.dgc_free(self.ctx.handle, _vp(self.ptr))
            self.ptr = 0


class Discretization:
    """Device copy of the geometry of `mesh` at order `order`.

    setup="host" computes the geometric factors and connectivity on the host
    (C++) and uploads them; setup="device" uploads only the cell corners and
    vertex ids and builds the same arrays, bit for bit, on the GPU (the
    choice for meshes of millions of cells)."""

    SETUPS = ("host", "device")

    def __init__(self, ctx: Context, mesh: Mesh, order: int, geometry: dict | None = None,
                 setup: str = "host"):
        if setup not in self.SETUPS:
            raise ValueError(f"setup must be one of {self.SETUPS}")
        self.ctx = ctx
        self.mesh = mesh
        self.order = order
        self.K = mesh.K
        self.Np = (order + 1) * (order + 2) // 2
        self.Nfp = order + 1
        self.setup = setup
        if setup == "device":
            if geometry is not None:
                raise ValueError("setup='device' builds the geometry itself")
            n = 3 * self.K
            cx, cy, ev = np.zeros(n), np.zeros(n), np.zeros(n, np.int32)
            check(lib().dgc_mesh_cells(mesh._h, _d(cx), _d(cy), _i(ev), None, None))
            h = _vp()
            check(lib().dgc_disc2d_create_from_cells(ctx.handle, self.K, mesh.Nv, order, _d(cx), _d(cy), _i(ev),
                                                     C.byref(h)))
            self._h = h
            self._geometry = None
            return
        g = geometry if geometry is not None else mesh.geometry(order)
        self._geometry = g
        ref = ref_operators(order, "weak")
        fm = ref["fmask"]
        f0 = np.ascontiguousarray(fm[0], np.int32)
        f1 = np.ascontiguousarray(fm[1], np.int32)
        f2 = np.ascontiguousarray(fm[2], np.int32)
        bx = np.ascontiguousarray(g["bx"]) if g["num_boundary_nodes"] else np.zeros(1)
        by = np.ascontiguousarray(g["by"]) if g["num_boundary_nodes"] else np.zeros(1)
        h = _vp()
        check(lib().dgc_disc2d_create(ctx.handle, self.K, order, _i(f0), _i(f1), _i(f2),
                                      _d(g["inv_jac"]), _d(g["J"]), _d(g["face_J"]),
                                      _i(g["iface_cell"]), _i(g["iface_face"]),
                                      g["num_boundary_nodes"], _d(bx), _d(by), _d(g["nx"]),
                                      _d(g["ny"]), C.byref(h)))
        self._h = h

    @property
    def geometry(self) -> dict:
        """Host geometry arrays (dgc_mesh_geometry layout); computed on first
        use for a device set-up."""
        if self._geometry is None:
            self._geometry = self.mesh.geometry(self.order)
        return self._geometry

    def arrays(self) -> dict:
        """Host copies of the device arrays: geo [13, K], nbr [3, K], J [K],
        face_J [3, K], bx, by."""
        K = self.K
        geo = np.zeros((13, K))
        nbr = np.zeros((3, K), np.int32)
        J = np.zeros(K)
        fJ = np.zeros((3, K))
        nb = C.c_int()
        check(lib().dgc_disc2d_arrays(self._h, _d(geo), _i(nbr), _d(J), _d(fJ), C.byref(nb), None, None))
        bx = np.zeros(max(nb.value, 1))
        by = np.zeros(max(nb.value, 1))
        check(lib().dgc_disc2d_arrays(self._h, None, None, None, None, None, _d(bx), _d(by)))
        return dict(geo=geo, nbr=nbr, J=J, face_J=fJ, bx=bx[:nb.value], by=by[:nb.value],
                    num_boundary_nodes=nb.value)

    @property
    def handle(self):
        return self._h

    @property
    def size(self):
        return self.K * self.Np

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None and self.ctx._h is not None:
            _LIB.dgc_disc2d_destroy(self._h)
            self._h = None


class _Operator:
    F = 1

    def __init__(self, disc: Discretization):
        self.disc = disc
        self.ctx = disc.ctx

    def new_state(self) -> list:
        return [DeviceArray(self.ctx, self.disc.size) for _ in range(self.F)]

    def new_work(self) -> list:
        return [DeviceArray(self.ctx, self.disc.size) for _ in range(5 * self.F)]

    def _fp(self, arrs):
        if len(arrs) != self.F:
            raise ValueError(f"expected {self.F} field arrays")
        return _ptrs([a.ptr for a in arrs])

    def rhs(self, inp, t, out):
        check(lib().dgc_op_rhs(self._h, self._fp(inp), self.disc.size, t, self._fp(out)))

    def rk4_step(self, state, t, dt, work):
        check(lib().dgc_rk4_step(self._h, self._fp(state), self.disc.size, t, dt,
                                 _ptrs([w.ptr for w in work])))

    def rk4_run(self, state, t0, t_end, dt, work):
        tout = C.c_double()
        steps = C.c_int()
        check(lib().dgc_rk4_run(self._h, self._fp(state), self.disc.size, t0, t_end, dt,
                                _ptrs([w.ptr for w in work]), C.byref(tout), C.byref(steps)))
        return tout.value, steps.value

    def rk4_steps(self, state, t0, dt, nsteps, work):
        check(lib().dgc_rk4_steps(self._h, self._fp(state), self.disc.size, t0, dt, nsteps,
                                  _ptrs([w.ptr for w in work])))

    def rk4_steps_timed(self, state, t0, dt, nsteps, work) -> float:
        ms = C.c_float()
        check(lib().dgc_rk4_steps_timed(self._h, self._fp(state), self.disc.size, t0, dt, nsteps,
                                        _ptrs([w.ptr for w in work]), C.byref(ms)))
        return ms.value

    @property
    def waves_per_tile(self) -> int:
        return lib().dgc_op_waves_per_tile(self._h)

    @waves_per_tile.setter
    def waves_per_tile(self, wt: int):
        check(lib().dgc_op_set_waves_per_tile(self._h, wt))

    @property
    def precision(self) -> str:
        """'fp64' or 'fp32-mfma' (FP32 matrix-core contractions, FP64 state)."""
        return "fp32-mfma" if lib().dgc_op_precision(self._h) == 1 else "fp64"

    @precision.setter
    def precision(self, value: str):
        check(lib().dgc_op_set_precision(self._h, {"fp64": 0, "fp32-mfma": 1}[value]))

    CACHE_POLICIES = {"auto": 0, "temporal": 1, "streaming": 2}

    @property
    def streaming(self) -> bool:
        """True if the stage kernels run with nontemporal accumulator streams."""
        return lib().dgc_op_cache_streaming(self._h) == 1

    def set_cache_policy(self, policy: str):
        """'auto' (stream when a step's working set exceeds the 256 MB
        last-level cache), 'temporal' or 'streaming'."""
        check(lib().dgc_op_set_cache_policy(self._h, self.CACHE_POLICIES[policy]))

    def autotune(self, state, work, reps=20) -> dict:
        """Time every built tiling on these buffers and keep the fastest (0 is the
        FP64 matrix-core variant)."""
        ms = (C.c_float * 4)()
        check(lib().dgc_op_autotune(self._h, self._fp(state), self.disc.size, _ptrs([w.ptr for w in work]),
                                    reps, ms))
        out = dict(choice=self.waves_per_tile, ms={wt: m for wt, m in zip((1, 2, 4, 0), ms) if m >= 0})
        if out["choice"] not in out["ms"]:  # 8 or 16 waves per tile, or the split tiling (32)
            out["ms"][out["choice"]] = self.stage_time(state, work, 2, reps)
        return out

    def stage_time(self, state, work, stage=2, reps=20) -> float:
        """Mean milliseconds of back-to-back launches of one stage kernel at the current tiling."""
        ms = C.c_float()
        check(lib().dgc_op_stage_time(self._h, self._fp(state), self.disc.size, _ptrs([w.ptr for w in work]),
                                      stage, reps, C.byref(ms)))
        return ms.value

    def profile_stages(self, state, t0, dt, nsteps, work):
        """Mean duration (ms) of each of the 4 stage kernels over nsteps steps."""
        ms = (C.c_float * 4)()
        check(lib().dgc_rk4_profile_stages(self._h, self._fp(state), self.disc.size, t0, dt, nsteps,
                                           _ptrs([w.ptr for w in work]), ms))
        return list(ms)

    def __del__(self):
        # a closed context has already released the device; skip the call
        if getattr(self, "_h", None) is not None and _LIB is not None and self.ctx._h is not None:
            _LIB.dgc_op_destroy(self._h)
            self._h = None


class MaxwellTM(_Operator):
    """Maxwell TM (Hx, Hy, Ez) with upwind flux and PEC walls."""

    F = 3

    def __init__(self, disc: Discretization, ops: dict | None = None):
        super().__init__(disc)
        ops = ops if ops is not None else ref_operators(disc.order, "weak")
        self.ops = ops
        h = _vp()
        check(lib().dgc_maxwell_tm_create(disc.handle, _d(np.ascontiguousarray(ops["Dr"])),
                                          _d(np.ascontiguousarray(ops["Ds"])),
                                          _d(np.ascontiguousarray(ops["LIFT"])), C.byref(h)))
        self._h = h


class Advection(_Operator):
    """Linear advection u_t + a.grad u = 0 with upwind flux."""

    F = 1
    BC_SINSIN = 0
    BC_ZERO = 1

    def __init__(self, disc: Discretization, ax=1.0, ay=1.0, bc=0, ops: dict | None = None):
        super().__init__(disc)
        ops = ops if ops is not None else ref_operators(disc.order, "weak")
        self.ops = ops
        h = _vp()
        check(lib().dgc_advection_create(disc.handle, _d(np.ascontiguousarray(ops["Dr"])),
                                         _d(np.ascontiguousarray(ops["Ds"])),
                                         _d(np.ascontiguousarray(ops["LIFT"])), ax, ay, bc,
                                         C.byref(h)))
        self._h = h
