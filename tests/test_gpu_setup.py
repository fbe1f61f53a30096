"""Device set-up of a discretization (dgc_disc2d_create_from_cells) against
the host set-up (dgc_mesh_geometry + dgc_disc2d_create): every device array
bit for bit, the same RK4 and IPDG results, and the host path's errors for
invalid cells. The host set-up itself is pinned to the reference's
geometry and connectivity by tests/test_mesh.py and the oracle tests."""
import ctypes as C
import os

import numpy as np
import pytest

import dgc_amd
from dgc_amd.core import lib, _d, _i, _vp

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _meshes():
    return {
        "bounded": lambda: dgc_amd.Mesh.structured(7, 5, 0.0, 1.0, 0.0, 0.7, jitter=0.2, seed=3),
        "periodic": lambda: dgc_amd.Mesh.structured(8, 6, -1.0, 1.0, -1.0, 1.0, jitter=0.2, seed=4, periodic=True),
        "gmsh_coarse": lambda: dgc_amd.Mesh.read_gmsh(os.path.join(GOLDEN, "square_domain_coarse.msh")),
        "gmsh_fine": lambda: dgc_amd.Mesh.read_gmsh(os.path.join(GOLDEN, "square_domain_fine.msh")),
    }


def _same(a, b):
    """Bitwise equality of two float arrays (NaN-free)."""
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def _compare(ctx, mesh, N):
    host = dgc_amd.Discretization(ctx, mesh, N)
    dev = dgc_amd.Discretization(ctx, mesh, N, setup="device")
    a, b = host.arrays(), dev.arrays()
    assert a["num_boundary_nodes"] == b["num_boundary_nodes"]
    assert np.array_equal(a["nbr"], b["nbr"])
    for k in ("geo", "J", "face_J", "bx", "by"):
        assert _same(np.ascontiguousarray(a[k]), np.ascontiguousarray(b[k])), k
    return host, dev


@pytest.mark.parametrize("name", list(_meshes()))
@pytest.mark.parametrize("N", [1, 3, 7])
def test_device_setup_matches_host_bit_for_bit(ctx, name, N):
    mesh = _meshes()[name]()
    _compare(ctx, mesh, N)


def test_device_setup_large_mesh(ctx):
    """Half a million cells: multi-block sort and scan of 1.5 M half-edges."""
    mesh = dgc_amd.Mesh.structured(512, 512, 0.0, 1.0, 0.0, 1.0, jitter=0.1, seed=5)
    _compare(ctx, mesh, 2)
    pm = dgc_amd.Mesh.structured(300, 200, 0.0, 1.0, 0.0, 1.0, jitter=0.1, seed=6, periodic=True)
    _compare(ctx, pm, 2)


def test_device_setup_gives_identical_solutions(ctx):
    mesh = dgc_amd.Mesh.structured(9, 7, -1.0, 1.0, -1.0, 1.0, jitter=0.15, seed=7)
    N = 3
    host, dev = _compare(ctx, mesh, N)
    x, y = mesh.dof_coords(N, node_major=True)
    u0 = np.sin(np.pi * x) * np.sin(np.pi * y)
    dt = 0.5 * mesh.min_edge / (N + 1) ** 2
    out = []
    for disc in (host, dev):
        # advection with the analytic inflow data at the boundary nodes
        op = dgc_amd.Advection(disc, 1.0, 1.0, dgc_amd.Advection.BC_SINSIN)
        st = [dgc_amd.DeviceArray.from_host(ctx, u0)]
        op.rk4_steps(st, 0.0, dt, 10, op.new_work())
        mx = dgc_amd.MaxwellTM(disc)
        sm = [dgc_amd.DeviceArray(ctx, disc.size), dgc_amd.DeviceArray(ctx, disc.size),
              dgc_amd.DeviceArray.from_host(ctx, u0)]
        sm[0].zero()
        sm[1].zero()
        mx.rk4_steps(sm, 0.0, dt, 10, mx.new_work())
        # IPDG apply (penalties from the host copies of J and face J)
        ip = dgc_amd.IPDG(disc, tau_mode=1)
        v = dgc_amd.DeviceArray.from_host(ctx, u0)
        ip.apply(v)
        out.append([st[0].download()] + [a.download() for a in sm] + [v.download()])
    for a, b in zip(*out):
        assert _same(a, b)


def _create(ctx, cx, cy, etov, nv, N=2):
    h = _vp()
    K = len(etov) // 3
    cx = np.ascontiguousarray(cx, np.float64)
    cy = np.ascontiguousarray(cy, np.float64)
    etov = np.ascontiguousarray(etov, np.int32)
    rc = lib().dgc_disc2d_create_from_cells(ctx.handle, K, nv, N, _d(cx), _d(cy), _i(etov), C.byref(h))
    if rc == 0:
        lib().dgc_disc2d_destroy(h)
    return rc, lib().dgc_last_error().decode()


def _two_cells():
    # unit square split along its diagonal, counter-clockwise
    vx = np.array([0.0, 1.0, 1.0, 0.0])
    vy = np.array([0.0, 0.0, 1.0, 1.0])
    etov = np.array([0, 1, 2, 0, 2, 3], np.int32)
    return vx, vy, etov


def test_device_setup_errors(ctx):
    vx, vy, etov = _two_cells()
    rc, msg = _create(ctx, vx[etov], vy[etov], etov, 4)
    assert rc == 0, msg
    # clockwise second cell
    cw = etov.copy()
    cw[4], cw[5] = cw[5], cw[4]
    rc, msg = _create(ctx, vx[cw], vy[cw], cw, 4)
    assert rc == 1 and "cell 1 is clockwise" in msg
    # vertex id out of range
    bad = etov.copy()
    bad[5] = 4
    rc, msg = _create(ctx, vx[etov], vy[etov], bad, 4)
    assert rc == 1 and "out of range" in msg
    # degenerate edge (a repeated vertex id)
    deg = etov.copy()
    deg[1] = 0
    rc, msg = _create(ctx, vx[etov], vy[etov], deg, 4)
    assert rc == 1 and "degenerate edge" in msg
    # three cells on one edge
    vx3 = np.array([0.0, 1.0, 1.0, 0.0, 2.0])
    vy3 = np.array([0.0, 0.0, 1.0, 1.0, 0.5])
    e3 = np.array([0, 1, 2, 0, 2, 3, 0, 4, 2], np.int32)
    rc, msg = _create(ctx, vx3[e3], vy3[e3], e3, 5)
    assert rc == 1 and "non-manifold" in msg
