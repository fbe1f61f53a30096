"""laplace_2d with Dirichlet or Neumann boundary faces and inhomogeneous data
on the GPU (dgc_ipdg_set_boundary, dgc_ipdg_apply, dgc_ipdg_apply_inhomogeneous)
against the compiled reference's golden vectors and the oracle (FP64, relative
L2 <= 1e-12), and a block-Jacobi PCG solve of the singular pure-Neumann problem."""
import numpy as np
import pytest

import dgc_amd
from dgc_amd import cell_to_node_major as c2n, node_to_cell_major as n2c
from oracle_lib import Problem

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


class Case:
    def __init__(self, ctx, mesh, N, tau_mode=0):
        self.ctx = ctx
        self.disc = dgc_amd.Discretization(ctx, mesh, N)
        self.ip = dgc_amd.IPDG(self.disc, tau_mode=tau_mode)
        self.K, self.Np = self.disc.K, self.disc.Np
        a = self.disc.arrays()
        bx, by = a["bx"], a["by"]
        self.gu = 2 * np.sin(bx) * np.cos(by)
        self.hq = 2 * (np.cos(bx) * np.sin(by) + bx * by)

    def dev(self, v):
        return dgc_amd.DeviceArray.from_host(self.ctx, c2n(np.ascontiguousarray(v, np.float64), self.K, self.Np))

    def host(self, a):
        return n2c(a.download(), self.K, self.Np)

    def hom(self, v, kind):
        self.ip.set_boundary(kind)
        d = self.dev(v)
        self.ip.apply(d)
        return self.host(d)

    def inh(self, kind):
        self.ip.set_boundary(kind)
        d = self.dev(np.zeros(self.K * self.Np))
        g = dgc_amd.DeviceArray.from_host(self.ctx, self.gu) if len(self.gu) else None
        h = dgc_amd.DeviceArray.from_host(self.ctx, self.hq) if len(self.hq) else None
        self.ip.apply_inhomogeneous(d, g, h)
        return self.host(d)


@pytest.mark.parametrize("name", ["grid", "coarse"])
def test_boundary_kinds_match_reference_golden(ctx, golden, name):
    fix = golden("laplace_bc_ref.npz")
    mesh = dgc_amd.Mesh.from_arrays(fix[f"{name}_vx"], fix[f"{name}_vy"], fix[f"{name}_etov"])
    c = Case(ctx, mesh, int(fix[f"{name}_order"]))
    for kind in (0, 1):
        inh = c.inh(kind)
        assert rel(inh, fix[f"{name}_inh_k{kind}"]) < TOL
        for k in range(2):
            hom = c.hom(fix[f"{name}_vin"][k], kind)
            assert rel(hom, fix[f"{name}_hom_k{kind}"][k]) < TOL
            assert rel(hom - inh, fix[f"{name}_full_k{kind}"][k]) < TOL


@pytest.mark.parametrize("N", range(1, 8))
@pytest.mark.parametrize("tau_mode", [0, 1])
def test_boundary_kinds_match_oracle(ctx, N, tau_mode):
    mesh = dgc_amd.Mesh.structured(13, 11, -1, 1, -1, 1, jitter=0.15, seed=N)
    c = Case(ctx, mesh, N, tau_mode)
    P = Problem.from_mesh(mesh, N, form="strong")
    u = np.random.default_rng(200 + N).standard_normal(c.K * c.Np)
    zero = np.zeros(c.K * c.Np)
    for kind in (1, 0):
        assert rel(c.hom(u, kind), P.laplace_bc(u, kind, tau_mode)) < TOL
        want = P.laplace_inhomogeneous(zero, kind, c.gu, c.hq, tau_mode)
        assert rel(c.inh(kind), want) < TOL


def test_pure_neumann_solve_with_block_jacobi(ctx):
    """-M Lap_h p = b with Neumann faces everywhere: singular (constants), solved
    by block-Jacobi PCG for a right-hand side orthogonal to the constants; the
    true residual is checked with the operator itself."""
    N = 3
    mesh = dgc_amd.Mesh.structured(24, 24, -1, 1, -1, 1, jitter=0.1, seed=8)
    c = Case(ctx, mesh, N)
    c.ip.set_boundary(dgc_amd.IPDG.NEUMANN)
    x, y = mesh.dof_coords(N, node_major=True)
    f = dgc_amd.DeviceArray.from_host(ctx, np.cos(np.pi * x) * np.cos(np.pi * y) + 0.3 * x)
    b = dgc_amd.DeviceArray(ctx, c.disc.size)
    c.ip.mass(f, b)
    bh = b.download()
    bh -= bh.mean()
    b = dgc_amd.DeviceArray.from_host(ctx, bh)
    p = dgc_amd.DeviceArray(ctx, c.disc.size)
    p.zero()
    res = c.ip.pcg(b, p, rtol=1e-10, maxit=5000, precond=1)
    assert res["relres"] <= 1e-10
    Sp = dgc_amd.DeviceArray(ctx, c.disc.size)
    c.ip.apply_spd(p, Sp)
    assert np.linalg.norm(Sp.download() - bh) / np.linalg.norm(bh) < 1e-8
    # two-level PCG: the coarse matrix is singular with Neumann faces; the coarse
    # solve inverts it plus a rank-one term along the coarse constants
    c.ip.coarse_space(mesh, 6)
    p2 = dgc_amd.DeviceArray(ctx, c.disc.size)
    p2.zero()
    res2 = c.ip.pcg(b, p2, rtol=1e-10, maxit=5000, precond=2)
    assert res2["relres"] <= 1e-10 and res2["iters"] < res["iters"] / 2
    c.ip.apply_spd(p2, Sp)
    assert np.linalg.norm(Sp.download() - bh) / np.linalg.norm(bh) < 1e-8
    # the two solutions differ by a constant only
    d = p2.download() - p.download()
    assert np.abs(d - d.mean()).max() < 1e-6 * np.abs(p.download()).max()
