"""1D advection on the GPU (dgc_adv1d_rhs, dgc_adv1d_rk4_steps) against the
compiled reference's golden vectors and the oracle (FP64, relative L2 <= 1e-12),
and the example's own run against the exact solution."""
import time

import numpy as np
import pytest

import dgc_amd
from dgc_amd import cell_to_node_major as c2n, node_to_cell_major as n2c
from oracle_lib import adv1d_rhs

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("K,N", [(16, 3), (64, 6)])
@pytest.mark.parametrize("own_operators", [False, True])
def test_adv1d_matches_reference_golden(ctx, golden, K, N, own_operators):
    z = golden("adv1d_ref.npz")
    p = f"K{K}_N{N}_"
    D, L = (None, None) if own_operators else (z[p + "D"], z[p + "L"])
    op = dgc_amd.Advection1D(ctx, K, N, D, L)
    Np = N + 1
    u = dgc_amd.DeviceArray.from_host(ctx, c2n(z[p + "ic"], K, Np))
    r = dgc_amd.DeviceArray(ctx, K * Np)
    op.rhs(u, 0.3, r)
    assert rel(n2c(r.download(), K, Np), z[p + "rhs03"]) < 1e-12
    dt = 0.25 / N / N * (2 * np.pi / K) / (2 * np.pi)
    op.rk4_steps(u, 0.0, dt, int(z[p + "steps"]))
    assert rel(n2c(u.download(), K, Np), z[p + "final"]) < 1e-12


@pytest.mark.parametrize("N", [1, 4, 9, 12])
def test_adv1d_rhs_matches_oracle(ctx, N):
    K = 1000
    D, L, x = dgc_amd.Advection1D.operators(K, N)
    op = dgc_amd.Advection1D(ctx, K, N)
    v = np.random.default_rng(N).standard_normal(K * (N + 1))
    r = dgc_amd.DeviceArray(ctx, K * (N + 1))
    op.rhs(dgc_amd.DeviceArray.from_host(ctx, v), 0.7, r)
    want = adv1d_rhs(K, N, D, L, n2c(v, K, N + 1), 0.7)
    assert rel(n2c(r.download(), K, N + 1), want) < 1e-12


def test_adv1d_example_run(ctx):
    """examples/advection_1d/main.cpp:83-116: 8192 cells, order 6, 10000 steps
    of dt = 0.25 / N^2 h / a from sin(x); mean-square error against
    sin(x - 2 pi t) as the example prints it."""
    K, N, steps = 8192, 6, 10000
    D, L, x = dgc_amd.Advection1D.operators(K, N)
    op = dgc_amd.Advection1D(ctx, K, N)
    u = dgc_amd.DeviceArray.from_host(ctx, np.sin(x))
    dt = 0.25 / N / N * (2 * np.pi / K) / (2 * np.pi)
    ctx.sync()
    t0 = time.perf_counter()
    op.rk4_steps(u, 0.0, dt, steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    err = np.mean((u.download() - np.sin(x - 2 * np.pi * steps * dt)) ** 2)
    print(f"t = {steps * dt}, error norm = {err:.3e}, time used: {wall * 1e3:.1f} ms")
    assert err < 1e-16
