"""The reference's fine mesh (examples/square_domain_fine.msh, K = 1064,
imported through the gmsh reader) end to end on the GPU: Maxwell TM and
advection RK4 against the oracle (FP64 relative L2 <= 1e-12) and the IPDG
Poisson solve judged by the oracle's FP64 residual."""
import os

import numpy as np
import pytest

import dgc_amd
from dgc_amd import node_to_cell_major as n2c, cell_to_node_major as c2n
from oracle_lib import Problem

pytestmark = pytest.mark.gpu
TOL = 1e-12
MESH = os.path.join(os.path.dirname(__file__), "golden", "square_domain_fine.msh")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("N", [2, 4])
def test_maxwell_and_advection_on_the_fine_mesh(ctx, N):
    mesh = dgc_amd.Mesh.read_gmsh(MESH)
    assert mesh.K == 1064
    disc = dgc_amd.Discretization(ctx, mesh, N)
    K, Np = disc.K, disc.Np
    P = Problem.from_mesh(mesh, N)
    x, y = mesh.dof_coords(N, node_major=False)
    # Maxwell: cavity mode initial data, 20 steps
    u0 = np.concatenate([np.zeros(2 * K * Np), np.sin(np.pi * x) * np.sin(np.pi * y)])
    op = dgc_amd.MaxwellTM(disc)
    dt = 0.5 * mesh.min_edge / (N + 1) ** 2
    st = [dgc_amd.DeviceArray.from_host(ctx, c2n(f, K, Np)) for f in np.split(u0, 3)]
    op.rk4_steps(st, 0.0, dt, 20, op.new_work())
    got = np.concatenate([n2c(a.download(), K, Np) for a in st])
    want, _ = P.rk4(0, u0, 0.0, dt, nsteps=20)
    err_mx = rel(got, want)
    # advection, the reference problem (a = (1, 1), analytic inflow), 20 steps
    adv = dgc_amd.Advection(disc, 1.0, 1.0, dgc_amd.Advection.BC_SINSIN)
    v0 = np.sin(np.pi * (x + 1)) * np.sin(np.pi * (y + 1))
    dta = 0.5 * mesh.min_edge / (np.sqrt(2.0) * (N + 1) ** 2)
    sa = [dgc_amd.DeviceArray.from_host(ctx, c2n(v0, K, Np))]
    adv.rk4_steps(sa, 0.0, dta, 20, adv.new_work())
    wa, _ = P.rk4(1, v0, 0.0, dta, nsteps=20, ax=1.0, ay=1.0, bc=0)
    err_adv = rel(n2c(sa[0].download(), K, Np), wa)
    print(f"fine mesh N={N}: Maxwell rel L2 {err_mx:.1e}, advection {err_adv:.1e}")
    assert err_mx < TOL and err_adv < TOL


def test_poisson_on_the_fine_mesh(ctx):
    N = 3
    mesh = dgc_amd.Mesh.read_gmsh(MESH)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.IPDG(disc, tau_mode=1)
    K, Np = disc.K, disc.Np
    xg, yg = mesh.dof_coords(N, node_major=True)
    f, u_exact = dgc_amd.poisson_rhs(xg, yg, seed=3)
    b = dgc_amd.DeviceArray(ctx, disc.size)
    op.mass(dgc_amd.DeviceArray.from_host(ctx, f), b)
    x = dgc_amd.DeviceArray(ctx, disc.size)
    res = op.pcg(b, x, rtol=1e-10, precond=1)
    P = Problem.from_mesh(mesh, N, form="strong")
    lap = P.laplace(n2c(x.download(), K, Np), 1).reshape(K, Np)
    M = dgc_amd.ref_operators(N, "strong")["M"]
    Sx = -(disc.geometry["J"][:, None] * (lap @ M.T)).reshape(-1)
    bc = n2c(b.download(), K, Np)
    rr = np.linalg.norm(bc - Sx) / np.linalg.norm(bc)
    err = np.abs(x.download() - u_exact).max()
    print(f"fine mesh Poisson N=3: {res['iters']} it, oracle relres {rr:.1e}, max error {err:.1e}")
    assert res["relres"] <= 1e-10 and rr <= 2e-10
    assert err < 1e-3
