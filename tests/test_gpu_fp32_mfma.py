"""FP32 matrix-core form of the RK stage (Dr u, Ds u and LIFT flux in FP32 on
16x16x4 MFMA; state, metric terms, fluxes and the RK4 update in FP64), for
the advection and Maxwell TM operators, against the oracle (the reference's
FP64 CPU path).

Stated tolerance of the FP32-contraction form: relative L2 <= 1e-5 against
the FP64 oracle for one RHS evaluation and for 20 RK4 steps, every order
N = 1..7 (the contraction operands are rounded to FP32, unit roundoff 6e-8,
and the derivative matrices amplify it by up to their norm)."""
import numpy as np
import pytest

import dgc_amd
from dgc_amd import node_to_cell_major as n2c, cell_to_node_major as c2n
from oracle_lib import Problem

pytestmark = pytest.mark.gpu
TOL_FP32 = 1e-5


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _setup(ctx, N, nside=24):
    mesh = dgc_amd.Mesh.structured(nside, nside, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.Advection(disc, 1.0, 0.5, dgc_amd.Advection.BC_ZERO)
    return mesh, disc, op


@pytest.mark.parametrize("N", range(1, 8))
def test_rhs_vs_oracle(ctx, N):
    mesh, disc, op = _setup(ctx, N)
    K, Np = disc.K, disc.Np
    x, y = mesh.dof_coords(N, node_major=False)
    u = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / 0.01) + 0.1 * np.random.default_rng(N).standard_normal(K * Np)
    want = Problem.from_mesh(mesh, N).advection_rhs(u, 0.0, ax=1.0, ay=0.5, bc=1)
    din = [dgc_amd.DeviceArray.from_host(ctx, c2n(u, K, Np))]
    dout = [dgc_amd.DeviceArray(ctx, K * Np)]
    op.precision = "fp32-mfma"
    assert op.precision == "fp32-mfma"
    op.rhs(din, 0.0, dout)
    err = rel(n2c(dout[0].download(), K, Np), want)
    op.precision = "fp64"
    op.rhs(din, 0.0, dout)
    err64 = rel(n2c(dout[0].download(), K, Np), want)
    print(f"N={N}: rhs rel L2 fp32-mfma {err:.2e}, fp64 {err64:.2e}")
    assert err < TOL_FP32
    assert err > 1e-12  # the FP32 form really ran
    assert err64 < 1e-13


@pytest.mark.parametrize("N", range(1, 8))
def test_rk4_vs_oracle(ctx, N):
    mesh, disc, op = _setup(ctx, N)
    K, Np = disc.K, disc.Np
    x, y = mesh.dof_coords(N, node_major=False)
    u0 = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / 0.01)
    dt = 0.5 * mesh.min_edge / (np.hypot(1.0, 0.5) * (N + 1) ** 2)
    op.precision = "fp32-mfma"
    state = [dgc_amd.DeviceArray.from_host(ctx, c2n(u0, K, Np))]
    op.rk4_steps(state, 0.0, dt, 20, op.new_work())
    want, _ = Problem.from_mesh(mesh, N).rk4(1, u0, 0.0, dt, nsteps=20, ax=1.0, ay=0.5, bc=1)
    err = rel(n2c(state[0].download(), K, Np), want)
    print(f"N={N}: 20 RK4 steps rel L2 fp32-mfma vs oracle {err:.2e}")
    assert err < TOL_FP32


def test_inflow_boundary_and_tail_tile(ctx, golden):
    """Time-dependent inflow data and K = 70 (not a multiple of the 64-cell tile)."""
    fix = golden("advection_ref.npz")
    N = int(fix["order"])
    mesh = dgc_amd.Mesh.from_arrays(fix["vx"], fix["vy"], fix["etov"])
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.Advection(disc, 1.0, 1.0, dgc_amd.Advection.BC_SINSIN)
    op.precision = "fp32-mfma"
    K, Np = disc.K, disc.Np
    state = [dgc_amd.DeviceArray.from_host(ctx, c2n(fix["ic"], K, Np))]
    op.rk4_steps(state, 0.0, float(fix["dt"]), int(fix["nsteps"]), op.new_work())
    assert rel(n2c(state[0].download(), K, Np), fix["final"]) < TOL_FP32
    mesh = dgc_amd.Mesh.structured(7, 5, -1, 1, -1, 1, jitter=0.1, seed=8)  # K = 70
    for N in (2, 5):
        disc = dgc_amd.Discretization(ctx, mesh, N)
        op = dgc_amd.Advection(disc, 1.0, 0.5, dgc_amd.Advection.BC_ZERO)
        K, Np = disc.K, disc.Np
        u = np.random.default_rng(3).standard_normal(K * Np)
        want = Problem.from_mesh(mesh, N).advection_rhs(u, 0.0, ax=1.0, ay=0.5, bc=1)
        op.precision = "fp32-mfma"
        din = [dgc_amd.DeviceArray.from_host(ctx, c2n(u, K, Np))]
        dout = [dgc_amd.DeviceArray(ctx, K * Np)]
        op.rhs(din, 0.0, dout)
        assert rel(n2c(dout[0].download(), K, Np), want) < TOL_FP32


@pytest.mark.parametrize("N", range(1, 8))
def test_maxwell_rhs_and_rk4_vs_oracle(ctx, N):
    """Maxwell TM (three fields, PEC walls, upwind flux) in the FP32 form: one
    RHS and 10 RK4 steps against the oracle, same stated tolerance."""
    mesh = dgc_amd.Mesh.structured(10, 10, -1, 1, -1, 1, jitter=0.1, seed=6)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.MaxwellTM(disc)
    op.precision = "fp32-mfma"
    assert op.precision == "fp32-mfma"
    K, Np = disc.K, disc.Np
    u0 = np.random.default_rng(N).standard_normal(3 * K * Np)
    P = Problem.from_mesh(mesh, N)
    up = lambda u: [dgc_amd.DeviceArray.from_host(ctx, c2n(f, K, Np)) for f in np.split(u, 3)]
    down = lambda a: np.concatenate([n2c(x.download(), K, Np) for x in a])
    out = [dgc_amd.DeviceArray(ctx, K * Np) for _ in range(3)]
    op.rhs(up(u0), 0.0, out)
    err_rhs = rel(down(out), P.maxwell_rhs(u0, 0.0))
    dt = 0.5 * mesh.min_edge / (N + 1) ** 2
    state = up(u0)
    op.rk4_steps(state, 0.0, dt, 10, op.new_work())
    want, _ = P.rk4(0, u0, 0.0, dt, nsteps=10)
    err = rel(down(state), want)
    print(f"Maxwell N={N}: fp32-mfma rhs rel L2 {err_rhs:.2e}, 10 RK4 steps {err:.2e}")
    assert 1e-12 < err_rhs < TOL_FP32
    assert err < TOL_FP32


def test_maxwell_cavity_fp32_form(ctx, golden):
    """The Maxwell cavity (K = 512, N = 4, to t = 1) in the FP32 form against
    the reference's golden final state."""
    fix = golden("maxwell_c1.npz")
    N = int(fix["order"])
    mesh = dgc_amd.Mesh.from_arrays(fix["vx"], fix["vy"], fix["etov"])
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.MaxwellTM(disc)
    op.precision = "fp32-mfma"
    K, Np = disc.K, disc.Np
    n = K * Np
    u0 = np.concatenate([np.zeros(2 * n), np.sin(np.pi * fix["x"]) * np.sin(np.pi * fix["y"])])
    state = [dgc_amd.DeviceArray.from_host(ctx, c2n(f, K, Np)) for f in np.split(u0, 3)]
    t, steps = op.rk4_run(state, 0.0, 1.0, float(fix["dt"]), op.new_work())
    assert steps == int(fix["steps"])
    got = np.concatenate([n2c(a.download(), K, Np) for a in state])
    err = rel(got, fix["final"])
    print(f"cavity fp32-mfma: {steps} steps, rel L2 vs reference {err:.2e}")
    assert err < TOL_FP32
