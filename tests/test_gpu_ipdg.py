"""GPU parity of the IPDG operator and its solvers against the reference's
golden vectors and the oracle (FP64: relative L2 <= 1e-12 for the operator;
the solvers are judged by the oracle's FP64 true residual)."""
import numpy as np
import pytest
from scipy.linalg import block_diag

import dgc_amd
from dgc_amd import cell_to_node_major as c2n, node_to_cell_major as n2c
from oracle_lib import Problem

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _from_fixture(fix, prefix=""):
    return dgc_amd.Mesh.from_arrays(fix[prefix + "vx"], fix[prefix + "vy"], fix[prefix + "etov"])


def test_apply_matches_reference(ctx, golden):
    fix = golden("laplace_ref.npz")
    N = int(fix["order"])
    mesh = _from_fixture(fix)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.IPDG(disc, tau_mode=0)
    K, Np = disc.K, disc.Np
    for vin, vout in zip(fix["vin"], fix["vout"]):
        u = dgc_amd.DeviceArray.from_host(ctx, c2n(vin, K, Np))
        op.apply(u)
        assert rel(n2c(u.download(), K, Np), vout) < 1e-13


def test_dense_operator_matches_reference(ctx, golden):
    fix = golden("laplace_ref.npz")
    N = int(fix["coarse_order"])
    mesh = _from_fixture(fix, "coarse_")
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.IPDG(disc, tau_mode=0)
    K, Np = disc.K, disc.Np
    n = K * Np
    cols = []
    for e in np.eye(n):
        u = dgc_amd.DeviceArray.from_host(ctx, c2n(e, K, Np))
        op.apply(u)
        cols.append(n2c(u.download(), K, Np))
    A = np.stack(cols, axis=1)
    np.testing.assert_allclose(A, fix["coarse_A"], atol=1e-12 * np.abs(A).max(), rtol=0)


@pytest.mark.parametrize("N,tau_mode", [(1, 1), (3, 1), (5, 0), (7, 1)])
def test_spd_form_and_mass_vs_oracle(ctx, N, tau_mode):
    mesh = dgc_amd.Mesh.structured(9, 8, -1, 1, -1, 1, jitter=0.1, seed=N)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.IPDG(disc, tau_mode=tau_mode)
    K, Np = disc.K, disc.Np
    P = Problem.from_mesh(mesh, N, form="strong")
    v = np.random.default_rng(N).standard_normal(K * Np)
    lap = P.laplace(v, tau_mode).reshape(K, Np)
    M = dgc_amd.ref_operators(N, "strong")["M"]
    J = disc.geometry["J"]
    want = -(J[:, None] * (lap @ M.T)).reshape(-1)
    u = dgc_amd.DeviceArray.from_host(ctx, c2n(v, K, Np))
    out = dgc_amd.DeviceArray(ctx, K * Np)
    op.apply_spd(u, out)
    assert rel(n2c(out.download(), K, Np), want) < 1e-12
    op.mass(u, out)
    wm = (J[:, None] * (v.reshape(K, Np) @ M.T)).reshape(-1)
    assert rel(n2c(out.download(), K, Np), wm) < 1e-13
    # in-place Laplacian in the SIPG mode too
    u2 = dgc_amd.DeviceArray.from_host(ctx, c2n(v, K, Np))
    op.apply(u2)
    assert rel(n2c(u2.download(), K, Np), lap.reshape(-1)) < 1e-12


def _dense_spd(P, mesh, N, tau_mode):
    n = P.K * P.Np
    A = np.stack([P.laplace(e, tau_mode) for e in np.eye(n)], axis=1)
    M = dgc_amd.ref_operators(N, "strong")["M"]
    J = mesh.geometry(N)["J"]
    return -block_diag(*[Jc * M for Jc in J]) @ A


def test_block_jacobi_blocks_are_exact_diagonal_blocks(ctx):
    N = 2
    mesh = dgc_amd.Mesh.structured(5, 4, -1, 1, -1, 1, jitter=0.1, seed=2)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.IPDG(disc, tau_mode=1)
    ncol = op.block_jacobi()
    assert 4 <= ncol <= 20
    blocks = op.diag_blocks()
    P = Problem.from_mesh(mesh, N, form="strong")
    S = _dense_spd(P, mesh, N, 1)
    Np = disc.Np
    for c in range(disc.K):
        want = S[c * Np:(c + 1) * Np, c * Np:(c + 1) * Np]
        np.testing.assert_allclose(blocks[c], want, atol=1e-12 * np.abs(S).max(), rtol=0)


@pytest.mark.parametrize("precond", [0, 1])
def test_pcg_small_matches_direct_solve(ctx, precond):
    N = 2
    mesh = dgc_amd.Mesh.structured(6, 6, -1, 1, -1, 1, jitter=0.1, seed=5)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.IPDG(disc, tau_mode=1)
    K, Np = disc.K, disc.Np
    P = Problem.from_mesh(mesh, N, form="strong")
    S = _dense_spd(P, mesh, N, 1)
    b = np.random.default_rng(3).standard_normal(K * Np)
    xd = dgc_amd.DeviceArray(ctx, K * Np)
    res = op.pcg(dgc_amd.DeviceArray.from_host(ctx, c2n(b, K, Np)), xd, rtol=1e-12, maxit=5000,
                 precond=precond, history=True)
    x = n2c(xd.download(), K, Np)
    assert rel(S @ x, b) < 2e-12
    assert rel(x, np.linalg.solve(S, b)) < 1e-9
    assert res["relres"] <= 1e-12 and res["iters"] > 0
    assert np.all(np.isfinite(res["history"]))


def test_rebuilt_block_jacobi_invalidates_captured_solver_graphs(ctx):
    """A second dgc_ipdg_block_jacobi frees and re-uploads the inverses: the
    captured PCG graphs that held the old addresses must be rebuilt, so a
    solve after the rebuild gives the same iterate as the one before it."""
    N = 2
    mesh = dgc_amd.Mesh.structured(6, 6, -1, 1, -1, 1, jitter=0.1, seed=5)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.IPDG(disc, tau_mode=1)
    K, Np = disc.K, disc.Np
    b = dgc_amd.DeviceArray.from_host(ctx, np.random.default_rng(3).standard_normal(K * Np))
    x1, x2 = dgc_amd.DeviceArray(ctx, K * Np), dgc_amd.DeviceArray(ctx, K * Np)
    r1 = op.pcg(b, x1, rtol=1e-12, maxit=5000, precond=1)
    op.block_jacobi()
    spacer = dgc_amd.DeviceArray(ctx, K * Np * Np)  # moves later allocations elsewhere
    op.block_jacobi()
    r2 = op.pcg(b, x2, rtol=1e-12, maxit=5000, precond=1)
    del spacer
    assert r1["iters"] == r2["iters"]
    np.testing.assert_array_equal(x1.download(), x2.download())


def _textbook_pcg(S, b, precond, iters):
    """Residual norms of textbook PCG (x0 = 0) on a dense matrix."""
    x = np.zeros_like(b)
    r = b.copy()
    z = precond(r)
    p = z.copy()
    rz = r @ z
    hist = [np.sqrt(r @ r)]
    for _ in range(iters):
        q = S @ p
        alpha = rz / (p @ q)
        x += alpha * p
        r -= alpha * q
        z = precond(r)
        rz_new = r @ z
        hist.append(np.sqrt(r @ r))
        p = z + (rz_new / rz) * p
        rz = rz_new
    return np.array(hist)


@pytest.mark.parametrize("precond", [0, 1])
def test_pcg_iterates_follow_textbook_pcg(ctx, precond):
    """The fused iteration (the search direction formed inside the operator's
    first pass; alpha, beta and the stopping test finished by the last
    workgroup of each reduction) follows textbook PCG on the dense operator:
    the residual norms of the first 25 iterations agree, and two solves are
    bitwise identical."""
    N = 2
    mesh = dgc_amd.Mesh.structured(6, 6, -1, 1, -1, 1, jitter=0.1, seed=5)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.IPDG(disc, tau_mode=1)
    K, Np = disc.K, disc.Np
    S = _dense_spd(Problem.from_mesh(mesh, N, form="strong"), mesh, N, 1)
    b = np.random.default_rng(4).standard_normal(K * Np)
    bd = dgc_amd.DeviceArray.from_host(ctx, c2n(b, K, Np))
    runs = []
    for _ in range(2):
        xd = dgc_amd.DeviceArray(ctx, K * Np)
        res = op.pcg(bd, xd, rtol=1e-12, maxit=5000, precond=precond, history=True)
        runs.append((res, xd.download()))
    assert runs[0][0]["iters"] == runs[1][0]["iters"]
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    np.testing.assert_array_equal(runs[0][0]["history"], runs[1][0]["history"])
    if precond == 1:
        inv = [np.linalg.inv(B) for B in op.diag_blocks()]
        inv = [0.5 * (A + A.T) for A in inv]
        apply = lambda r: np.concatenate([inv[c] @ r[c * Np:(c + 1) * Np] for c in range(K)])
    else:
        apply = lambda r: r.copy()
    want = _textbook_pcg(S, b, apply, 25)
    got = runs[0][0]["history"][:26]
    d = np.max(np.abs(got - want) / want)
    print(f"precond {precond}: {runs[0][0]['iters']} it; max rel diff of the first 25 residual norms {d:.1e}")
    assert d < 1e-8


def _poisson_problem(ctx, n_side, N, seed=3):
    mesh = dgc_amd.Mesh.structured(n_side, n_side, -1, 1, -1, 1, jitter=0.1, seed=seed)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.IPDG(disc, tau_mode=1)
    x, y = mesh.dof_coords(N, node_major=True)
    f, u_exact = dgc_amd.poisson_rhs(x, y, seed=seed)
    fd = dgc_amd.DeviceArray.from_host(ctx, f)
    b = dgc_amd.DeviceArray(ctx, disc.size)
    op.mass(fd, b)
    return mesh, disc, op, b, u_exact


def _oracle_relres(mesh, disc, x_node, b_node):
    """FP64 true residual ||b - S x|| / ||b|| computed by the oracle."""
    N, K, Np = disc.order, disc.K, disc.Np
    P = Problem.from_mesh(mesh, N, form="strong")
    x = n2c(x_node, K, Np)
    lap = P.laplace(x, 1).reshape(K, Np)
    M = dgc_amd.ref_operators(N, "strong")["M"]
    Sx = -(disc.geometry["J"][:, None] * (lap @ M.T)).reshape(-1)
    b = n2c(b_node, K, Np)
    return np.linalg.norm(b - Sx) / np.linalg.norm(b)


def test_poisson_32k_n3_block_jacobi_pcg_to_1e10(ctx):
    """The Poisson benchmark problem: 128x128 squares (K = 32768), N = 3."""
    mesh, disc, op, b, u_exact = _poisson_problem(ctx, 128, 3)
    x = dgc_amd.DeviceArray(ctx, disc.size)
    ncol = op.block_jacobi()
    res = op.pcg(b, x, rtol=1e-10, maxit=20000, precond=1)
    xs = x.download()
    rr = _oracle_relres(mesh, disc, xs, b.download())
    err = np.abs(xs - u_exact).max()
    print(f"PCG(block-Jacobi) K={disc.K} N=3: {res['iters']} it, {res['ms']:.1f} ms, "
          f"relres {res['relres']:.2e} (oracle {rr:.2e}), max err {err:.2e}, {ncol} colours")
    assert res["relres"] <= 1e-10 and rr <= 2e-10
    assert err < 1e-4


def test_pcg_above_the_fused_reduction_limit(ctx):
    """Above 4096 partial sums (K > 262144 cells) the CG scalars are finished
    by separate one-workgroup kernels instead of inside the iteration's
    kernels: 363 x 363 squares (K = 263538), N = 1, block-Jacobi PCG to
    1e-10 by the oracle's FP64 residual."""
    mesh, disc, op, b, u_exact = _poisson_problem(ctx, 363, 1)
    assert (disc.K + 63) // 64 > 4096
    x = dgc_amd.DeviceArray(ctx, disc.size)
    res = op.pcg(b, x, rtol=1e-10, maxit=40000, precond=1, history=True)
    rr = _oracle_relres(mesh, disc, x.download(), b.download())
    print(f"PCG(block-Jacobi) K={disc.K} N=1: {res['iters']} it, {res['ms']:.1f} ms, oracle relres {rr:.2e}")
    assert res["relres"] <= 1e-10 and rr <= 2e-10
    h = res["history"]
    assert np.all(np.isfinite(h)) and h[-1] <= 1e-10 * np.linalg.norm(b.download())


def test_poisson_mixed_precision_refinement(ctx):
    mesh, disc, op, b, u_exact = _poisson_problem(ctx, 64, 3)
    x = dgc_amd.DeviceArray(ctx, disc.size)
    res = op.pcg_mixed(b, x, rtol=1e-10, inner_rtol=1e-4)
    xs = x.download()
    rr = _oracle_relres(mesh, disc, xs, b.download())
    x64 = dgc_amd.DeviceArray(ctx, disc.size)
    r64 = op.pcg(b, x64, rtol=1e-10, precond=1)
    print(f"mixed K={disc.K}: {res['outer']} outer, {res['inner_iters']} inner it, {res['ms']:.1f} ms; "
          f"FP64 PCG {r64['iters']} it {r64['ms']:.1f} ms; oracle relres {rr:.2e}")
    assert res["relres"] <= 1e-10 and rr <= 2e-10
    assert rel(xs, x64.download()) < 1e-8


@pytest.mark.parametrize("N", [2, 3, 5, 7])
def test_shifted_spd_form_vs_oracle(ctx, N):
    """dgc_ipdg_set_shift: the SPD form -J M Lap_h + sigma J M against the
    oracle's Laplacian and mass product; sigma = 0 restores the plain form bit
    for bit; the fused PCG (first chunk captured with the shift) solves the
    shifted system, with Dirichlet and with Neumann faces (nonsingular)."""
    mesh = dgc_amd.Mesh.structured(9, 8, -1, 1, -1, 1, jitter=0.1, seed=N)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.IPDG(disc, tau_mode=1)
    K, Np = disc.K, disc.Np
    P = Problem.from_mesh(mesh, N, form="strong")
    M = dgc_amd.ref_operators(N, "strong")["M"]
    J = disc.geometry["J"]
    v = np.random.default_rng(50 + N).standard_normal(K * Np)
    u = dgc_amd.DeviceArray.from_host(ctx, c2n(v, K, Np))
    out = dgc_amd.DeviceArray(ctx, K * Np)
    op.apply_spd(u, out)
    plain = out.download()
    sigma = 750.0
    for kind in (0, 1):
        op.set_boundary(kind)
        op.set_shift(sigma)
        op.apply_spd(u, out)
        lap = P.laplace_bc(v, kind, 1).reshape(K, Np)
        want = (J[:, None] * ((sigma * v.reshape(K, Np) - lap) @ M.T)).reshape(-1)
        assert rel(n2c(out.download(), K, Np), want) < 1e-12
        b = out.download()
        x = dgc_amd.DeviceArray(ctx, K * Np)
        x.zero()
        res = op.pcg(dgc_amd.DeviceArray.from_host(ctx, b), x, rtol=1e-11, maxit=2000, precond=1)
        assert res["relres"] <= 1e-11
        assert rel(x.download(), c2n(v, K, Np)) < 1e-6
        op.set_shift(0.0)
    op.set_boundary(0)
    op.apply_spd(u, out)
    assert np.array_equal(out.download(), plain)


def test_uncaptured_solver_launches_give_identical_solves(ctx):
    """dgc_ipdg_set_graphs(0) launches the iterations' kernels one by one (for
    kernel tracers) instead of replaying captured graphs: the same iterates,
    bit for bit, for block-Jacobi and two-level PCG."""
    N = 3
    mesh = dgc_amd.Mesh.structured(16, 16, -1, 1, -1, 1, jitter=0.1, seed=5)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.IPDG(disc, tau_mode=1)
    op.coarse_space(mesh, 4)
    b = dgc_amd.DeviceArray.from_host(ctx, np.random.default_rng(7).standard_normal(disc.size))
    for precond in (1, 2):
        xs = []
        for graphs in (True, False):
            op.set_graphs(graphs)
            x = dgc_amd.DeviceArray(ctx, disc.size)
            r = op.pcg(b, x, rtol=1e-10, maxit=5000, precond=precond)
            xs.append((r["iters"], x.download()))
        op.set_graphs(True)
        assert xs[0][0] == xs[1][0]
        np.testing.assert_array_equal(xs[0][1], xs[1][1])
