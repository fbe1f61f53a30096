"""The Stokes dual-splitting step on the GPU (dgc_stokes_advance) against its
CPU restatement (oracle/stokes_oracle.py) and the example's exact solution, and
the BR1 operator with boundary data (dgc_ipdg_apply_data) against the oracle."""
import os
import sys

import numpy as np
import pytest

import dgc_amd
from dgc_amd import cell_to_node_major as c2n, node_to_cell_major as n2c
from oracle_lib import Problem

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import stokes_oracle as so  # noqa: E402

pytestmark = pytest.mark.gpu
FINE = os.path.join(os.path.dirname(__file__), "golden", "square_domain_fine.msh")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("N", range(1, 8))
def test_laplace_with_boundary_data_matches_oracle(ctx, N):
    mesh = dgc_amd.Mesh.structured(11, 9, -1, 1, -1, 1, jitter=0.15, seed=N)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    ip = dgc_amd.IPDG(disc)
    P = Problem.from_mesh(mesh, N, form="strong")
    K, Np = disc.K, disc.Np
    a = disc.arrays()
    rng = np.random.default_rng(300 + N)
    u = rng.standard_normal(K * Np)
    g = rng.standard_normal(len(a["bx"]))
    for kind in (0, 1):
        ip.set_boundary(kind)
        d = dgc_amd.DeviceArray.from_host(ctx, c2n(u, K, Np))
        gd = dgc_amd.DeviceArray.from_host(ctx, g)
        if kind == 0:
            ip.apply_data(d, gd=gd)
            want = P.laplace_data(u, 0, g, None)
        else:
            ip.apply_data(d, hd=gd)
            want = P.laplace_data(u, 1, None, g)
        assert rel(n2c(d.download(), K, Np), want) < 1e-12


def _run_gpu(ctx, mesh, N, dt, steps, rtol=1e-13, coarse=None):
    disc = dgc_amd.Discretization(ctx, mesh, N)
    st = dgc_amd.Stokes(disc)
    if coarse:
        st.coarse_space(mesh, coarse)
    x, y = mesh.dof_coords(N, node_major=True)
    ux, uy = so.exact(x, y, 0.0)
    up = (dgc_amd.DeviceArray.from_host(ctx, ux), dgc_amd.DeviceArray.from_host(ctx, uy))
    u = (dgc_amd.DeviceArray.from_host(ctx, ux), dgc_amd.DeviceArray.from_host(ctx, uy))
    nxt = (dgc_amd.DeviceArray(ctx, disc.size), dgc_amd.DeviceArray(ctx, disc.size))
    tp = t = 0.0
    its = []
    for _ in range(steps):
        its.append(st.advance(up, tp, u, t, dt, nxt, rtol))
        up, u, nxt = u, nxt, up
        tp, t = t, t + dt
    return disc, u[0].download(), u[1].download(), t, its


def test_stokes_step_matches_oracle(ctx):
    mesh = dgc_amd.Mesh.structured(6, 6, -1, 1, -1, 1, jitter=0.1, seed=3)
    N, dt, steps = 2, 1e-3, 3
    disc, gx, gy, t, its = _run_gpu(ctx, mesh, N, dt, steps)
    K, Np = disc.K, disc.Np
    ow, os_ = dgc_amd.ref_operators(N, "weak"), dgc_amd.ref_operators(N, "strong")
    Pw, Ps = Problem.from_mesh(mesh, N, ops=ow), Problem.from_mesh(mesh, N, ops=os_)
    c = mesh.cells()
    geo = so.Geometry(c["cx"], c["cy"], c["nb_cell"], os_["r"], os_["s"], os_["fmask"])
    x, y = mesh.dof_coords(N, node_major=False)
    up = u = so.exact(x, y, 0.0)
    tp = tc = 0.0
    for _ in range(steps):
        un = so.advance_timestep(Pw, Ps, geo, os_["M"], up, tp, u, tc, dt)
        up, tp, u, tc = u, tc, un, tc + dt
    print("iterations per step:", its)
    assert rel(n2c(gx, K, Np), u[0]) < 1e-9 and rel(n2c(gy, K, Np), u[1]) < 1e-9


def test_stokes_example_on_the_fine_mesh(ctx):
    """The example's configuration (examples/stokes_2d/main.cpp:65-196): the
    reference's fine mesh (K = 1064), dt = 1e-3 to T = 0.1, u^{n-1} = u^n at the
    start; mean-square error against the exact solution, falling with the order."""
    mesh = dgc_amd.Mesh.read_gmsh(FINE)
    errs = []
    for N in (1, 2):
        disc, gx, gy, t, its = _run_gpu(ctx, mesh, N, 1e-3, 100, rtol=1e-11)
        x, y = mesh.dof_coords(N, node_major=True)
        ex, ey = so.exact(x, y, t)
        errs.append((np.sum((gx - ex) ** 2) + np.sum((gy - ey) ** 2)) / len(x))
        print(f"N={N}: T={t:.3f} error norm {errs[-1]:.3e}, last step iterations {its[-1]}")
    assert abs(t - 0.1) < 1e-12
    assert errs[0] < 1e-3 and errs[1] < errs[0] / 5


def test_stokes_with_two_level_pressure_solves(ctx):
    """The same steps with the two-level pressure preconditioner: the same
    velocities (to the solver tolerance) in fewer pressure iterations."""
    mesh = dgc_amd.Mesh.structured(32, 32, -1, 1, -1, 1, jitter=0.1, seed=3)
    _, ax, ay, _, its1 = _run_gpu(ctx, mesh, 2, 1e-3, 3, rtol=1e-12)
    _, bx, by, _, its2 = _run_gpu(ctx, mesh, 2, 1e-3, 3, rtol=1e-12, coarse=8)
    print("block-Jacobi:", its1, "two-level:", its2)
    assert rel(bx, ax) < 1e-8 and rel(by, ay) < 1e-8
    assert its2[-1]["pressure"] < its1[-1]["pressure"] / 2



def test_stokes_reset_forgets_the_retained_pressure(ctx):
    """Each pressure solve starts from the handle's last pressure; after
    dgc_stokes_reset the first step of a new trajectory repeats the fresh
    handle's first step exactly."""
    mesh = dgc_amd.Mesh.structured(6, 6, -1, 1, -1, 1, jitter=0.1, seed=3)
    N, dt = 2, 1e-3
    disc = dgc_amd.Discretization(ctx, mesh, N)
    st = dgc_amd.Stokes(disc)
    x, y = mesh.dof_coords(N, node_major=True)
    ux, uy = so.exact(x, y, 0.0)
    u = (dgc_amd.DeviceArray.from_host(ctx, ux), dgc_amd.DeviceArray.from_host(ctx, uy))
    out = (dgc_amd.DeviceArray(ctx, disc.size), dgc_amd.DeviceArray(ctx, disc.size))
    st.advance(u, 0.0, u, 0.0, dt, out, 1e-13)
    first = (out[0].download(), out[1].download())
    for _ in range(2):
        st.advance(u, 0.0, u, 0.0, dt, out, 1e-13)
    st.reset()
    st.advance(u, 0.0, u, 0.0, dt, out, 1e-13)
    np.testing.assert_array_equal(out[0].download(), first[0])
    np.testing.assert_array_equal(out[1].download(), first[1])
