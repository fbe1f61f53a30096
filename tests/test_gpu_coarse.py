"""Two-level preconditioner for the IPDG system on the GPU: the coarse
Galerkin matrix against Phi^T S Phi formed with the oracle's operator, the
basis functions against the DOF coordinates, the FP32 inverse, and two-level
PCG judged by the oracle's FP64 residual."""
import numpy as np
import pytest

import dgc_amd
from oracle_lib import Problem

pytestmark = pytest.mark.gpu


def _oracle_S(mesh, N):
    """Dense -J M Lap_h (cell-major) from the oracle's apply, column by column."""
    P = Problem.from_mesh(mesh, N, form="strong")
    K, Np = P.K, P.Np
    M = dgc_amd.ref_operators(N, "strong")["M"]
    J = mesh.geometry(N)["J"]
    n = K * Np
    S = np.zeros((n, n))
    for i in range(n):
        u = np.zeros(n)
        u[i] = 1.0
        lap = P.laplace(u, 1).reshape(K, Np)
        S[:, i] = (-(J[:, None] * (lap @ M.T))).reshape(-1)
    return S


def _phi(info, K, Np, nc):
    """Phi (cell-major DOFs x coarse unknowns) from the device's coefficients."""
    Phi = np.zeros((K * Np, nc))
    r, s = info["r"], info["s"]
    cf = info["coef"].astype(np.float64)
    for c in range(K):
        a = info["cell_agg"][c]
        rows = c * Np + np.arange(Np)
        Phi[rows, 3 * a] = 1.0
        Phi[rows, 3 * a + 1] = cf[0, c] + cf[1, c] * r + cf[2, c] * s
        Phi[rows, 3 * a + 2] = cf[3, c] + cf[4, c] * r + cf[5, c] * s
    return Phi


def test_coarse_matrix_is_the_galerkin_product(ctx):
    N, nside, g = 3, 8, 4
    mesh = dgc_amd.Mesh.structured(nside, nside, -1, 1, -1, 1, jitter=0.1, seed=3)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.IPDG(disc, tau_mode=1)
    nc = op.coarse_space(mesh, g)
    assert nc == 3 * g * g
    info = op.coarse_info()
    K, Np = disc.K, disc.Np
    Phi = _phi(info, K, Np, nc)
    # the basis is (1, (x - x_a)/h, (y - y_a)/h) on each box of the 4 x 4 grid
    x, y = mesh.dof_coords(N, node_major=False)
    cells = mesh.cells()
    x0, x1, y0, y1 = cells["cx"].min(), cells["cx"].max(), cells["cy"].min(), cells["cy"].max()
    wx, wy = (x1 - x0) / g, (y1 - y0) / g
    h = 0.5 * max(wx, wy)
    cxm = cells["cx"].reshape(-1, 3).mean(1)
    cym = cells["cy"].reshape(-1, 3).mean(1)
    box = (np.minimum((cxm - x0) // wx, g - 1) + g * np.minimum((cym - y0) // wy, g - 1)).astype(int)
    assert np.array_equal(info["cell_agg"], box)  # every box holds cells: ids are box ids
    cell_of_dof = np.repeat(np.arange(K), Np)
    agg = box[cell_of_dof]
    xa = x0 + (agg % g + 0.5) * wx
    ya = y0 + (agg // g + 0.5) * wy
    cols = np.arange(K * Np)
    assert np.max(np.abs(Phi[cols, 3 * agg + 1] * h + xa - x)) < 1e-6
    assert np.max(np.abs(Phi[cols, 3 * agg + 2] * h + ya - y)) < 1e-6
    # Galerkin matrix against the oracle's operator
    S = _oracle_S(mesh, N)
    want = Phi.T @ S @ Phi
    A = info["A"]
    err = np.max(np.abs(A - want)) / np.max(np.abs(want))
    print(f"coarse matrix rel err {err:.2e}")
    assert err < 1e-12
    assert np.all(np.linalg.eigvalsh(0.5 * (A + A.T)) > 0)
    E = info["Ainv"].astype(np.float64) @ A - np.eye(nc)
    print(f"|Ainv A - I|_max {np.max(np.abs(E)):.2e}")
    assert np.max(np.abs(E)) < 1e-3
    assert np.array_equal(info["Ainv"], info["Ainv"].T)


def _poisson(ctx, n_side, N, seed=3):
    mesh = dgc_amd.Mesh.structured(n_side, n_side, -1, 1, -1, 1, jitter=0.1, seed=seed)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.IPDG(disc, tau_mode=1)
    x, y = mesh.dof_coords(N, node_major=True)
    f, _ = dgc_amd.poisson_rhs(x, y, seed=seed)
    b = dgc_amd.DeviceArray(ctx, disc.size)
    op.mass(dgc_amd.DeviceArray.from_host(ctx, f), b)
    return mesh, disc, op, b


def _oracle_relres(mesh, disc, x_node, b_node):
    N, K, Np = disc.order, disc.K, disc.Np
    P = Problem.from_mesh(mesh, N, form="strong")
    lap = P.laplace(dgc_amd.node_to_cell_major(x_node, K, Np), 1).reshape(K, Np)
    M = dgc_amd.ref_operators(N, "strong")["M"]
    Sx = -(disc.geometry["J"][:, None] * (lap @ M.T)).reshape(-1)
    b = dgc_amd.node_to_cell_major(b_node, K, Np)
    return np.linalg.norm(b - Sx) / np.linalg.norm(b)


def test_two_level_pcg_iterates_follow_textbook_pcg(ctx):
    """The fused two-level iteration (the prolongation's last workgroup
    finishes beta and the stopping test) follows textbook PCG with
    z = P_BJ r + Phi Ainv Phi^T r on the oracle's dense operator, Ainv the
    device's FP32 inverse; two solves are bitwise identical."""
    N, nside, g = 2, 8, 2
    mesh = dgc_amd.Mesh.structured(nside, nside, -1, 1, -1, 1, jitter=0.1, seed=7)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.IPDG(disc, tau_mode=1)
    K, Np = disc.K, disc.Np
    nc = op.coarse_space(mesh, g)
    b = np.random.default_rng(6).standard_normal(K * Np)
    bd = dgc_amd.DeviceArray.from_host(ctx, dgc_amd.cell_to_node_major(b, K, Np))
    runs = []
    for _ in range(2):
        xd = dgc_amd.DeviceArray(ctx, K * Np)
        res = op.pcg(bd, xd, rtol=1e-12, maxit=5000, precond=2, history=True)
        runs.append((res, xd.download()))
    assert runs[0][0]["iters"] == runs[1][0]["iters"]
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    info = op.coarse_info()
    Phi = _phi(info, K, Np, nc)
    Ainv = info["Ainv"].astype(np.float64)
    inv = [np.linalg.inv(B) for B in op.diag_blocks()]
    inv = [0.5 * (A + A.T) for A in inv]

    def apply(r):
        zb = np.concatenate([inv[c] @ r[c * Np:(c + 1) * Np] for c in range(K)])
        return zb + Phi @ (Ainv @ (Phi.T @ r))

    S = _oracle_S(mesh, N)
    x = np.zeros_like(b)
    r = b.copy()
    z = apply(r)
    p = z.copy()
    rz = r @ z
    want = [np.sqrt(r @ r)]
    for _ in range(20):
        q = S @ p
        alpha = rz / (p @ q)
        x += alpha * p
        r -= alpha * q
        z = apply(r)
        rz_new = r @ z
        want.append(np.sqrt(r @ r))
        p = z + (rz_new / rz) * p
        rz = rz_new
    got = runs[0][0]["history"][:21]
    d = np.max(np.abs(got - np.array(want)) / np.array(want))
    print(f"two-level: {runs[0][0]['iters']} it; max rel diff of the first 20 residual norms {d:.1e}")
    assert d < 1e-8


def test_two_level_pcg_converges_in_fewer_iterations(ctx):
    mesh, disc, op, b = _poisson(ctx, 32, 3)
    xb = dgc_amd.DeviceArray(ctx, disc.size)
    rb = op.pcg(b, xb, rtol=1e-10, precond=1)
    op.coarse_space(mesh, 8)
    x2 = dgc_amd.DeviceArray(ctx, disc.size)
    r2 = op.pcg(b, x2, rtol=1e-10, precond=2, history=True)
    rr = _oracle_relres(mesh, disc, x2.download(), b.download())
    print(f"block-Jacobi {rb['iters']} it {rb['ms']:.1f} ms; two-level {r2['iters']} it {r2['ms']:.1f} ms; "
          f"oracle relres {rr:.2e}")
    assert r2["relres"] <= 1e-10 and rr <= 2e-10
    assert r2["iters"] <= 0.5 * rb["iters"]
    # same solution as block-Jacobi PCG to the solver tolerance
    d = np.linalg.norm(x2.download() - xb.download()) / np.linalg.norm(xb.download())
    assert d < 1e-8
    # rebuilding the coarse space (new arrays) keeps the solver consistent
    op.coarse_space(mesh, 4)
    x3 = dgc_amd.DeviceArray(ctx, disc.size)
    r3 = op.pcg(b, x3, rtol=1e-10, precond=2)
    assert r3["relres"] <= 1e-10
    assert _oracle_relres(mesh, disc, x3.download(), b.download()) <= 2e-10
