"""Batched ensemble of Maxwell TM instances on the GPU: one launch per order
and RK4 stage for instances of mixed orders and sizes must give each
instance exactly what its own per-order kernels give, and the reference's
CPU path through the oracle (FP64, relative L2 <= 1e-12)."""
import numpy as np
import pytest

import dgc_amd
from dgc_amd import node_to_cell_major as n2c, cell_to_node_major as c2n
from oracle_lib import Problem

pytestmark = pytest.mark.gpu

SPECS = [(2, 6, 6), (3, 9, 7), (4, 12, 12), (5, 5, 5), (6, 8, 8), (2, 20, 20), (1, 7, 5), (7, 4, 4)]


def _instances(ctx, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for i, (N, nx, ny) in enumerate(SPECS):
        mesh = dgc_amd.Mesh.structured(nx, ny, -1, 1, -1, 1, jitter=0.1, seed=10 + i)
        disc = dgc_amd.Discretization(ctx, mesh, N)
        op = dgc_amd.MaxwellTM(disc)
        u0 = rng.standard_normal(3 * disc.K * disc.Np)
        dt = 0.5 * mesh.min_edge / (N + 1) ** 2
        out.append((mesh, disc, op, u0, dt))
    return out


def _upload(ctx, u, K, Np):
    return [dgc_amd.DeviceArray.from_host(ctx, c2n(f, K, Np)) for f in np.split(u, 3)]


def _download(arrs, K, Np):
    return np.concatenate([n2c(a.download(), K, Np) for a in arrs])


def test_ensemble_equals_per_instance_kernels_and_oracle(ctx):
    inst = _instances(ctx)
    states = [_upload(ctx, u0, d.K, d.Np) for (_, d, _, u0, _) in inst]
    works = [op.new_work() for (_, _, op, _, _) in inst]
    ens = dgc_amd.Ensemble(ctx, [i[2] for i in inst], states, works, [i[4] for i in inst])
    info = ens.info()
    print(f"ensemble: {info}")
    assert info["num_items"] >= len(inst) and info["grid"] >= 1
    ens.steps(10)
    for k, (mesh, disc, op, u0, dt) in enumerate(inst):
        got = _download(states[k], disc.K, disc.Np)
        # the per-instance kernels at the tiling the ensemble runs (the small
        # instances' own default, 16 waves per tile, has no ensemble kernel)
        op.waves_per_tile = info["waves_per_tile"][k]
        ref_state = _upload(ctx, u0, disc.K, disc.Np)
        op.rk4_steps(ref_state, 0.0, dt, 10, op.new_work())
        want = _download(ref_state, disc.K, disc.Np)
        d = np.max(np.abs(got - want)) / np.max(np.abs(want))
        print(f"instance {k} N={disc.order} K={disc.K}: max rel diff vs per-instance kernels {d:.1e}")
        assert d == 0.0  # same device code, same tiling: identical arithmetic
        if k in (1, 7):
            oracle, _ = Problem.from_mesh(mesh, disc.order).rk4(0, u0, 0.0, dt, nsteps=10)
            assert np.linalg.norm(got - oracle) / np.linalg.norm(oracle) < 1e-12


def test_matrix_core_tiling_in_the_ensemble(ctx):
    """Instances on the FP64 matrix-core variant (tiling 0) give what their own
    per-order matrix-core kernels give."""
    inst = _instances(ctx, seed=2)
    for _, _, op, _, _ in inst:
        op.waves_per_tile = 0
    states = [_upload(ctx, u0, d.K, d.Np) for (_, d, _, u0, _) in inst]
    works = [op.new_work() for (_, _, op, _, _) in inst]
    ens = dgc_amd.Ensemble(ctx, [i[2] for i in inst], states, works, [i[4] for i in inst])
    assert ens.info()["waves_per_tile"] == [0] * len(inst)
    ens.steps(5)
    for k, (mesh, disc, op, u0, dt) in enumerate(inst):
        ref_state = _upload(ctx, u0, disc.K, disc.Np)
        op.rk4_steps(ref_state, 0.0, dt, 5, op.new_work())
        got = _download(states[k], disc.K, disc.Np)
        want = _download(ref_state, disc.K, disc.Np)
        assert np.max(np.abs(got - want)) == 0.0, k


@pytest.mark.parametrize("tiling", ["default", "mfma"])
def test_streaming_kernels_in_the_ensemble(ctx, tiling):
    """The ensemble's nontemporal-stream kernels give bitwise what its
    temporal kernels give (the small ensemble runs temporal under the
    automatic policy)."""
    inst = _instances(ctx, seed=3)
    if tiling == "mfma":
        for _, _, op, _, _ in inst:
            op.waves_per_tile = 0
    states = [_upload(ctx, u0, d.K, d.Np) for (_, d, _, u0, _) in inst]
    works = [op.new_work() for (_, _, op, _, _) in inst]
    ens = dgc_amd.Ensemble(ctx, [i[2] for i in inst], states, works, [i[4] for i in inst])
    assert not ens.streaming
    ens.steps(4)
    temporal = [_download(s, d.K, d.Np) for s, (_, d, _, _, _) in zip(states, inst)]
    for s, (_, d, _, u0, _) in zip(states, inst):
        for a, f in zip(s, np.split(u0, 3)):
            a.upload(c2n(f, d.K, d.Np))
    ens.set_cache_policy("streaming")
    assert ens.streaming
    ens.steps(4)
    for k, (s, (_, d, _, _, _)) in enumerate(zip(states, inst)):
        np.testing.assert_array_equal(_download(s, d.K, d.Np), temporal[k], err_msg=str(k))


def test_autotuned_and_sequential_modes_agree(ctx):
    inst = _instances(ctx, seed=1)
    states = [_upload(ctx, u0, d.K, d.Np) for (_, d, _, u0, _) in inst]
    works = [op.new_work() for (_, _, op, _, _) in inst]
    ens = dgc_amd.Ensemble(ctx, [i[2] for i in inst], states, works, [i[4] for i in inst])
    ens.autotune(reps=5)
    print(f"autotuned: {ens.info()['waves_per_tile']}")
    ens.steps(6)
    got = [_download(s, d.K, d.Np) for s, (_, d, _, _, _) in zip(states, inst)]
    # the same steps again from the initial data, instance by instance
    for s, (_, d, _, u0, _) in zip(states, inst):
        for a, f in zip(s, np.split(u0, 3)):
            a.upload(c2n(f, d.K, d.Np))
    ens.steps_sequential(6)
    for k, (s, (_, d, _, _, _)) in enumerate(zip(states, inst)):
        seq = _download(s, d.K, d.Np)
        assert np.linalg.norm(seq - got[k]) / np.linalg.norm(seq) < 1e-13, k


def test_rejects_non_maxwell_operator(ctx):
    mesh = dgc_amd.Mesh.structured(4, 4, -1, 1, -1, 1)
    disc = dgc_amd.Discretization(ctx, mesh, 2)
    op = dgc_amd.Advection(disc, 1.0, 0.5, dgc_amd.Advection.BC_ZERO)
    st = [dgc_amd.DeviceArray(ctx, disc.size) for _ in range(3)]
    with pytest.raises(dgc_amd.DGCError):
        dgc_amd.Ensemble(ctx, [op], [st], [[dgc_amd.DeviceArray(ctx, disc.size) for _ in range(12)]], [1e-3])
