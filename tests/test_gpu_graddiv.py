"""grad_2d / div_2d on the GPU (dgc_grad2d_apply, dgc_div2d_apply) against the
compiled reference's golden vectors and the oracle (FP64, relative L2 <= 1e-12),
and the closed form the oracle is pinned by (exact gradients of polynomials)."""
import os

import numpy as np
import pytest

import dgc_amd
from dgc_amd import cell_to_node_major as c2n, node_to_cell_major as n2c
from oracle_lib import Problem

pytestmark = pytest.mark.gpu
TOL = 1e-12
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


class Case:
    """A discretization, its GPU operator and helpers in the CPU (cell-major) layout."""

    def __init__(self, ctx, mesh, N):
        self.ctx, self.mesh, self.N = ctx, mesh, N
        self.disc = dgc_amd.Discretization(ctx, mesh, N)
        self.op = dgc_amd.GradDiv(self.disc)
        self.K, self.Np = self.disc.K, self.disc.Np
        a = self.disc.arrays()
        self.bx, self.by = a["bx"], a["by"]

    def up(self, v):
        return dgc_amd.DeviceArray.from_host(self.ctx, c2n(np.ascontiguousarray(v, np.float64), self.K, self.Np))

    def gup(self, g):
        return None if g is None else dgc_amd.DeviceArray.from_host(self.ctx, np.ascontiguousarray(g))

    def grad(self, u, g=None, beta=1.0):
        ox, oy = dgc_amd.DeviceArray(self.ctx, self.K * self.Np), dgc_amd.DeviceArray(self.ctx, self.K * self.Np)
        self.op.grad(self.up(u), ox, oy, self.gup(g), beta)
        return n2c(ox.download(), self.K, self.Np), n2c(oy.download(), self.K, self.Np)

    def div(self, u0, u1, g0=None, g1=None, beta=1.0):
        out = dgc_amd.DeviceArray(self.ctx, self.K * self.Np)
        self.op.div(self.up(u0), self.up(u1), out, self.gup(g0), self.gup(g1), beta)
        return n2c(out.download(), self.K, self.Np)

    def bcs(self):
        bx, by = self.bx, self.by
        return {0: (None, None, 1.0), 1: (2 * np.sin(bx) * np.cos(by), 2 * np.cos(bx) * np.sin(by), -1.0)}


@pytest.mark.parametrize("name", ["coarse", "grid"])
def test_grad_div_match_reference_golden(ctx, golden, name):
    fix = golden("graddiv_ref.npz")
    mesh = dgc_amd.Mesh.from_arrays(fix[f"{name}_vx"], fix[f"{name}_vy"], fix[f"{name}_etov"])
    c = Case(ctx, mesh, int(fix[f"{name}_order"]))
    for bc, (g0, g1, beta) in c.bcs().items():
        for k in range(2):
            u0, u1 = fix[f"{name}_vin"][k]
            gx, gy = c.grad(u0, g0, beta)
            want = fix[f"{name}_grad_bc{bc}"][k]
            assert rel(gx, want[0]) < TOL and rel(gy, want[1]) < TOL
            assert rel(c.div(u0, u1, g0, g1, beta), fix[f"{name}_div_bc{bc}"][k]) < TOL


@pytest.mark.parametrize("N", range(1, 8))
@pytest.mark.parametrize("periodic", [False, True])
def test_grad_div_match_oracle(ctx, N, periodic):
    mesh = dgc_amd.Mesh.structured(23, 17, -1, 1, -1, 1, jitter=0.15, seed=N, periodic=periodic)
    c = Case(ctx, mesh, N)
    P = Problem.from_mesh(mesh, N)
    rng = np.random.default_rng(100 + N)
    u0, u1 = rng.standard_normal((2, c.K * c.Np))
    for bc, (g0, g1, beta) in c.bcs().items():
        if periodic and bc == 1:
            continue  # no boundary faces
        gx, gy = c.grad(u0, g0, beta)
        wx, wy = P.grad(u0, g0, beta)
        assert rel(gx, wx) < TOL and rel(gy, wy) < TOL
        assert rel(c.div(u0, u1, g0, g1, beta), P.div(u0, u1, g0, g1, beta)) < TOL


@pytest.mark.parametrize("N", [4, 7])
def test_grad_div_exact_for_polynomials(ctx, N):
    mesh = dgc_amd.Mesh.structured(9, 7, -1, 1, -1, 1, jitter=0.15, seed=5)
    c = Case(ctx, mesh, N)
    x, y = mesh.dof_coords(N, node_major=False)
    bx, by = c.bx, c.by
    p = lambda x, y: x**N + 2 * x * y**(N - 1) - 3 * y**N + 0.5  # noqa: E731
    q = lambda x, y: y**N - x * y**(N - 1)  # noqa: E731
    px = N * x**(N - 1) + 2 * y**(N - 1)
    py = 2 * (N - 1) * x * y**(N - 2) - 3 * N * y**(N - 1)
    qy = N * y**(N - 1) - (N - 1) * x * y**(N - 2)
    gx, gy = c.grad(p(x, y), 2 * p(bx, by), -1.0)
    np.testing.assert_allclose(gx, px, atol=1e-9)
    np.testing.assert_allclose(gy, py, atol=1e-9)
    d = c.div(p(x, y), q(x, y), 2 * p(bx, by), 2 * q(bx, by), -1.0)
    np.testing.assert_allclose(d, px + qy, atol=1e-9)


def test_grad_div_reject_aliasing(ctx):
    mesh = dgc_amd.Mesh.structured(4, 4)
    c = Case(ctx, mesh, 2)
    u = c.up(np.zeros(c.K * c.Np))
    v = dgc_amd.DeviceArray(ctx, c.K * c.Np)
    with pytest.raises(dgc_amd.DGCError, match="alias"):
        c.op.grad(u, u, v)
    with pytest.raises(dgc_amd.DGCError, match="alias"):
        c.op.div(u, v, v)


@pytest.mark.parametrize("N", [1, 2, 3, 5])
def test_grad_div_tilings_agree(ctx, N):
    mesh = dgc_amd.Mesh.structured(40, 30, -1, 1, -1, 1, jitter=0.15, seed=N)
    c = Case(ctx, mesh, N)
    rng = np.random.default_rng(400 + N)
    u0, u1 = rng.standard_normal((2, c.K * c.Np))
    g0, g1, beta = c.bcs()[1]
    base = c.grad(u0, g0, beta) + (c.div(u0, u1, g0, g1, beta),)
    for wt in (1, 2, 4):
        c.op.set_tiling(wt)
        got = c.grad(u0, g0, beta) + (c.div(u0, u1, g0, g1, beta),)
        for a, b in zip(got, base):
            assert rel(a, b) < 1e-14

