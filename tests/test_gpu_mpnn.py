"""The learned preconditioner on the GPU: forward and gradients of the
project's kernels against the float64 oracle (which mirrors the kernels'
bf16 rounding points), GPU training, and flexible CG judged by the oracle's
FP64 residual against block-Jacobi PCG and against the untrained network."""
import os
import sys

import numpy as np
import pytest

import dgc_amd
from conftest import ROOT
from oracle_lib import Problem

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mpnn_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


def _flatten(p, L):
    return np.concatenate([p[k].reshape(-1) for k in
                           ("A1", "B1", "b1", "A2", "B2", "b2", "A3", "B3", "b3", "Wo", "bo")]).astype(np.float32)


def _setup(ctx, n_side=6, N=3, seed=0):
    mesh = dgc_amd.Mesh.structured(n_side, n_side, -1, 1, -1, 1, jitter=0.1, seed=seed)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.IPDG(disc, tau_mode=1)
    net = dgc_amd.LearnedPreconditioner(op, seed=1)
    c = mesh.cells()
    nbr = c["nb_cell"].reshape(-1, 3).astype(int)
    nbr[nbr == np.arange(mesh.K)[:, None]] = -1
    return mesh, disc, op, net, nbr


def _random_model(net, nbr, seed=3):
    rng = np.random.default_rng(seed)
    p = mo.init_params(net.D0, 64, net.Np, seed)
    for k in p:
        if k.startswith("b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.1
    X = rng.standard_normal((net.K, net.D0))
    X[:, 2 * net.Np + 4:] = 0.0
    return p, X


def test_forward_matches_oracle(ctx):
    mesh, disc, op, net, nbr = _setup(ctx)
    p, X = _random_model(net, nbr)
    net.set_params(_flatten(p, None))
    y = net.eval(X.astype(np.float32))
    want, _ = mo.forward(p, X.astype(np.float32), nbr, bf=True)
    err = np.linalg.norm(y - want) / np.linalg.norm(want)
    print(f"forward rel err {err:.2e}")
    assert err < 1e-6


def test_gradients_match_oracle(ctx):
    mesh, disc, op, net, nbr = _setup(ctx)
    p, X = _random_model(net, nbr, seed=4)
    net.set_params(_flatten(p, None))
    T = np.random.default_rng(9).standard_normal((net.K, net.Np)).astype(np.float32)
    loss, g = net.loss_grad(X.astype(np.float32), T)
    wl, wg, _ = mo.loss_and_grads(p, X.astype(np.float32), T, nbr, bf=True)
    assert abs(loss - wl) <= 1e-6 * wl
    off = 0
    for k in ("A1", "B1", "b1", "A2", "B2", "b2", "A3", "B3", "b3", "Wo", "bo"):
        n = p[k].size
        gg, ww = g[off:off + n], wg[k].reshape(-1)
        off += n
        err = np.linalg.norm(gg - ww) / max(np.linalg.norm(ww), 1e-30)
        print(f"grad {k}: rel err {err:.2e}")
        assert err < 1e-5, k


def test_energy_loss_gradients_match_oracle(ctx):
    """Energy-norm training step (block-Jacobi step, features, network, S z on
    the device) against the oracle's energy loss with the same operator."""
    mesh, disc, op, net, nbr = _setup(ctx)
    K, Np = net.K, net.Np
    p, _ = _random_model(net, nbr, seed=5)
    net.set_params(_flatten(p, None))
    ct = 0.3
    net.target_scale = ct
    assert net.target_scale == ct

    def apply_S(v):
        a = dgc_amd.DeviceArray.from_host(ctx, dgc_amd.cell_to_node_major(np.reshape(v, -1), K, Np))
        out = dgc_amd.DeviceArray(ctx, disc.size)
        op.apply_spd(a, out)
        return dgc_amd.node_to_cell_major(out.download(), K, Np).reshape(K, Np)

    e = np.random.default_rng(11).standard_normal((K, Np))
    r = apply_S(e)
    er = float(np.sum(e * r))
    zb = np.linalg.solve(op.diag_blocks(), r[..., None])[..., 0]
    x, s_r, s_z = mo.features(r, zb, disc.geometry["inv_jac"].reshape(K, 4), Np)
    X = np.zeros((K, net.D0))
    X[:, :2 * Np + 4] = x
    wl, wg, wy = mo.energy_loss_and_grads(p, X, nbr, zb, r, e, s_z * ct, apply_S, bf=True)
    rd = dgc_amd.DeviceArray.from_host(ctx, dgc_amd.cell_to_node_major(r.reshape(-1), K, Np))
    zd = dgc_amd.DeviceArray(ctx, disc.size)
    loss, g = net.energy_loss_grad(rd, er, zd)
    z = dgc_amd.node_to_cell_major(zd.download(), K, Np).reshape(K, Np)
    dz = np.linalg.norm((z - zb) - s_z * ct * wy) / np.linalg.norm(s_z * ct * wy)
    print(f"energy loss {loss:.6f} oracle {wl:.6f}; correction rel err {dz:.2e}")
    assert dz < 1e-6
    assert abs(loss - wl) <= 1e-6 * abs(wl)
    off = 0
    for k in ("A1", "B1", "b1", "A2", "B2", "b2", "A3", "B3", "b3", "Wo", "bo"):
        n = p[k].size
        gg, ww = g[off:off + n], wg[k].reshape(-1)
        off += n
        err = np.linalg.norm(gg - ww) / max(np.linalg.norm(ww), 1e-30)
        print(f"energy grad {k}: rel err {err:.2e}")
        assert err < 1e-5, k


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


@pytest.mark.parametrize("base,N", [(1, 3), (2, 3), (2, 2)])
def test_train_and_flexible_cg(ctx, base, N):
    """Untrained (zero output layer) FCG reproduces PCG with the base
    preconditioner; training lowers the held-out energy loss; FCG with the
    trained network reaches the FP64 tolerance by the oracle's residual
    (N = 2 on the two-level base: the setting of the 131072-cell benchmark
    leg, here on a 32 x 32 mesh)."""
    mesh, disc, op, b = _poisson(ctx, 32, N)
    if base == 2:
        op.coarse_space(mesh, 8)
    net = dgc_amd.LearnedPreconditioner(op, seed=0, base=base)
    net.build_dataset(mesh, 96, seed0=100, modes=16)
    x0 = dgc_amd.DeviceArray(ctx, disc.size)
    r0 = net.fcg(b, x0, rtol=1e-10)
    xp = dgc_amd.DeviceArray(ctx, disc.size)
    rp = op.pcg(b, xp, rtol=1e-10, precond=base)
    before = net.eval_loss(64, 32)
    hist = net.train(0, 64, epochs=4, lr=3e-4, seed=0)
    after = net.eval_loss(64, 32)
    x1 = dgc_amd.DeviceArray(ctx, disc.size)
    r1 = net.fcg(b, x1, rtol=1e-10)
    rr = _oracle_relres(mesh, disc, x1.download(), b.download())
    print(f"base {base}: train loss {np.round(hist, 4)}, held-out {before:.4f} -> {after:.4f}; FCG untrained "
          f"{r0['iters']} it {r0['ms']:.1f} ms, trained {r1['iters']} it {r1['ms']:.1f} ms (fallback "
          f"{r1['fell_back']}); PCG {rp['iters']} it {rp['ms']:.1f} ms; oracle relres {rr:.2e}")
    assert abs(r0["iters"] - rp["iters"]) <= 2
    assert after < before
    assert r1["relres"] <= 1e-10 and rr <= 2e-10


def test_training_pairs_from_the_solver_itself(ctx):
    """dgc_mpnn_build_dataset_solver: pairs (r = S e, e = x* - x_j) from the
    recorded iterates of base-preconditioned PCG solves: e.r is positive, and
    the base preconditioner (the untrained network) leaves a larger share of
    the error energy on a late iterate (the hard modes) than on the first one
    (x_0 = 0, e = x*)."""
    mesh, disc, op, _ = _poisson(ctx, 32, 3)
    op.coarse_space(mesh, 8)
    net = dgc_amd.LearnedPreconditioner(op, seed=0, base=2)
    n = net.build_dataset_solver(mesh, 4, seed0=100, modes=16, snapshots=4)
    assert n == 16
    first = net.eval_loss(0, 1)
    late = net.eval_loss(3, 1)
    print(f"base energy ratio: first pair {first:.4f}, fourth pair {late:.4f}")
    assert 0 < first < late < 1.0
    hist = net.train(0, 12, epochs=2, lr=3e-4, seed=0)
    assert np.all(np.isfinite(hist))


def test_network_is_refused_after_an_operator_change(ctx):
    """The network's training data and target scale belong to the operator it
    was created for: after dgc_ipdg_set_shift (or set_boundary) flexible CG,
    training and evaluation return an error instead of running on a stale
    preconditioner."""
    mesh, disc, op, b = _poisson(ctx, 8, 2)
    net = dgc_amd.LearnedPreconditioner(op, seed=0, base=1)
    net.build_dataset(mesh, 4, seed0=100, modes=4)
    x = dgc_amd.DeviceArray(ctx, disc.size)
    assert net.fcg(b, x, rtol=1e-8)["relres"] <= 1e-8
    op.set_shift(2.0)
    with pytest.raises(dgc_amd.DGCError, match="operator changed"):
        net.fcg(b, x, rtol=1e-8)
    with pytest.raises(dgc_amd.DGCError, match="operator changed"):
        net.train(0, 4, epochs=1)
    op.set_shift(0.0)
    op.pcg(b, x, rtol=1e-8, precond=1)  # rebuilds the block inverses; the network stays stale
    with pytest.raises(dgc_amd.DGCError, match="operator changed"):
        net.eval_loss(0, 4)
