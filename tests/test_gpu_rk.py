"""GPU parity of the fused RHS + RK4 path (through the C ABI) against the
reference's golden vectors and the oracle, FP64: relative L2 <= 1e-12."""
import numpy as np
import pytest

import dgc_amd
from dgc_amd import node_to_cell_major as n2c, cell_to_node_major as c2n
from oracle_lib import Problem

pytestmark = pytest.mark.gpu
TOL = 1e-12  # FP64 parity bar: relative L2 against the reference CPU path


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def upload_fields(ctx, u_cell, F, K, Np):
    return [dgc_amd.DeviceArray.from_host(ctx, c2n(f, K, Np)) for f in np.split(u_cell, F)]


def download_fields(arrs, K, Np):
    return np.concatenate([n2c(a.download(), K, Np) for a in arrs])


@pytest.fixture(scope="module")
def c1(ctx, golden):
    fix = golden("maxwell_c1.npz")
    N = int(fix["order"])
    mesh = dgc_amd.Mesh.from_arrays(fix["vx"], fix["vy"], fix["etov"])
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.MaxwellTM(disc)
    return fix, mesh, disc, op


def test_maxwell_rhs_matches_reference_and_oracle(ctx, c1):
    fix, mesh, disc, op = c1
    K, Np = disc.K, disc.Np
    n = K * Np
    u = np.concatenate([np.zeros(2 * n), np.sin(np.pi * fix["x"]) * np.sin(np.pi * fix["y"])])
    rng = np.random.default_rng(1)
    for trial in range(2):
        if trial == 1:
            u = rng.standard_normal(3 * n)
        din = upload_fields(ctx, u, 3, K, Np)
        dout = [dgc_amd.DeviceArray(ctx, n) for _ in range(3)]
        op.rhs(din, 0.0, dout)
        got = download_fields(dout, K, Np)
        want = Problem.from_mesh(mesh, disc.order).maxwell_rhs(u)
        assert rel(got, want) < 1e-13, trial
        if trial == 0:
            assert rel(got, fix["rhs0"]) < 1e-13


def test_maxwell_cavity_t1_matches_reference(ctx, c1):
    """The headline parity case: K=512, N=4, t = 1 (491 RK4 steps)."""
    fix, mesh, disc, op = c1
    K, Np = disc.K, disc.Np
    n = K * Np
    u0 = np.concatenate([np.zeros(2 * n), np.sin(np.pi * fix["x"]) * np.sin(np.pi * fix["y"])])
    state = upload_fields(ctx, u0, 3, K, Np)
    work = op.new_work()
    t, steps = op.rk4_run(state, 0.0, 1.0, float(fix["dt"]), work)
    assert steps == int(fix["steps"]) and t == float(fix["t_final"])
    got = download_fields(state, K, Np)
    err = rel(got, fix["final"])
    print(f"maxwell C1 t=1 rel L2 vs reference = {err:.3e}")
    assert err < TOL
    x, y, s2 = fix["x"], fix["y"], np.sqrt(2.0)
    exact = np.concatenate([-np.sin(np.pi * x) * np.cos(np.pi * y) * np.sin(s2 * np.pi) / s2,
                            np.cos(np.pi * x) * np.sin(np.pi * y) * np.sin(s2 * np.pi) / s2,
                            np.sin(np.pi * x) * np.sin(np.pi * y) * np.cos(s2 * np.pi)])
    assert np.mean((got - exact) ** 2) < 1e-8


@pytest.mark.parametrize("wt", [4, 8, 16, 32])
def test_maxwell_cavity_t1_every_small_mesh_tiling(ctx, golden, wt):
    """The cavity run at 4, 8 and 16 waves per tile and with each tile's rows
    split over 4 workgroups (32, the default on this 8-tile mesh at N = 4):
    the golden t = 1 state to the FP64 bar."""
    fix = golden("maxwell_c1.npz")
    mesh = dgc_amd.Mesh.from_arrays(fix["vx"], fix["vy"], fix["etov"])
    disc = dgc_amd.Discretization(ctx, mesh, int(fix["order"]))
    op = dgc_amd.MaxwellTM(disc)
    assert op.waves_per_tile == 32
    op.waves_per_tile = wt
    K, Np = disc.K, disc.Np
    u0 = np.concatenate([np.zeros(2 * K * Np), np.sin(np.pi * fix["x"]) * np.sin(np.pi * fix["y"])])
    state = upload_fields(ctx, u0, 3, K, Np)
    t, steps = op.rk4_run(state, 0.0, 1.0, float(fix["dt"]), op.new_work())
    assert steps == int(fix["steps"]) and rel(download_fields(state, K, Np), fix["final"]) < TOL


def test_single_steps_equal_graph_replay(ctx, c1):
    fix, mesh, disc, op = c1
    K, Np = disc.K, disc.Np
    rng = np.random.default_rng(2)
    u0 = rng.standard_normal(3 * K * Np)
    a = upload_fields(ctx, u0, 3, K, Np)
    b = upload_fields(ctx, u0, 3, K, Np)
    wa, wb = op.new_work(), op.new_work()
    dt = float(fix["dt"])
    # 37 steps: two 16-step graph launches and five one-step launches
    for i in range(37):
        op.rk4_step(a, i * dt, dt, wa)
    op.rk4_steps(b, 0.0, dt, 37, wb)
    np.testing.assert_array_equal(download_fields(a, K, Np), download_fields(b, K, Np))
    # and one step against the reference's first step
    s = upload_fields(ctx, np.concatenate([np.zeros(2 * K * Np),
                                           np.sin(np.pi * fix["x"]) * np.sin(np.pi * fix["y"])]), 3, K, Np)
    op.rk4_step(s, 0.0, dt, wa)
    assert rel(download_fields(s, K, Np), fix["step1"]) < 1e-13


def test_advection_reference_problem(ctx, golden):
    """a = (1, 1) with the reference's analytic inflow data (time-dependent BC)."""
    fix = golden("advection_ref.npz")
    N = int(fix["order"])
    mesh = dgc_amd.Mesh.from_arrays(fix["vx"], fix["vy"], fix["etov"])
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.Advection(disc, 1.0, 1.0, dgc_amd.Advection.BC_SINSIN)
    K, Np = disc.K, disc.Np
    din = upload_fields(ctx, fix["ic"], 1, K, Np)
    dout = [dgc_amd.DeviceArray(ctx, K * Np)]
    op.rhs(din, float(fix["t_rhs"]), dout)
    assert rel(download_fields(dout, K, Np), fix["rhs"]) < 1e-13
    state = upload_fields(ctx, fix["ic"], 1, K, Np)
    work = op.new_work()
    op.rk4_steps(state, 0.0, float(fix["dt"]), int(fix["nsteps"]), work)
    assert rel(download_fields(state, K, Np), fix["final"]) < TOL
    # graph replays driven by the device clock equal explicit-time single steps
    s2 = upload_fields(ctx, fix["ic"], 1, K, Np)
    dt = float(fix["dt"])
    t = 0.0
    for _ in range(int(fix["nsteps"])):
        op.rk4_step(s2, t, dt, work)
        t += dt
    np.testing.assert_array_equal(download_fields(s2, K, Np), download_fields(state, K, Np))
    # 16-step graphs carry the device clock across their steps: 35 steps from
    # the time loop equal 35 explicit-time steps
    s3 = upload_fields(ctx, fix["ic"], 1, K, Np)
    s4 = upload_fields(ctx, fix["ic"], 1, K, Np)
    op.rk4_steps(s3, 0.25, dt, 35, work)
    t = 0.25
    for _ in range(35):
        op.rk4_step(s4, t, dt, work)
        t += dt
    np.testing.assert_array_equal(download_fields(s3, K, Np), download_fields(s4, K, Np))


@pytest.mark.parametrize("N", range(1, 8))
def test_periodic_advection_all_orders_vs_oracle(ctx, N):
    """Config-2 physics (a = (1, 0.5), periodic) on a 24x24 mesh, 20 steps."""
    mesh = dgc_amd.Mesh.structured(24, 24, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.Advection(disc, 1.0, 0.5, dgc_amd.Advection.BC_ZERO)
    K, Np = disc.K, disc.Np
    x, y = mesh.dof_coords(N, node_major=False)
    u0 = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / 0.01)
    dt = 0.5 * mesh.min_edge / (np.hypot(1.0, 0.5) * (N + 1) ** 2)
    state = upload_fields(ctx, u0, 1, K, Np)
    op.rk4_steps(state, 0.0, dt, 20, op.new_work())
    got = download_fields(state, K, Np)
    want, _ = Problem.from_mesh(mesh, N).rk4(1, u0, 0.0, dt, nsteps=20, ax=1.0, ay=0.5, bc=1)
    assert rel(got, want) < TOL


@pytest.mark.parametrize("N", range(1, 8))
def test_fp64_matrix_core_variant_vs_oracle(ctx, N):
    """The advection operator's FP64 MFMA variant (tiling 0) meets the FP64 bar."""
    mesh = dgc_amd.Mesh.structured(24, 24, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.Advection(disc, 1.0, 0.5, dgc_amd.Advection.BC_ZERO)
    op.waves_per_tile = 0
    assert op.waves_per_tile == 0
    K, Np = disc.K, disc.Np
    x, y = mesh.dof_coords(N, node_major=False)
    u0 = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / 0.01)
    dt = 0.5 * mesh.min_edge / (np.hypot(1.0, 0.5) * (N + 1) ** 2)
    state = upload_fields(ctx, u0, 1, K, Np)
    op.rk4_steps(state, 0.0, dt, 20, op.new_work())
    want, _ = Problem.from_mesh(mesh, N).rk4(1, u0, 0.0, dt, nsteps=20, ax=1.0, ay=0.5, bc=1)
    err = rel(download_fields(state, K, Np), want)
    print(f"N={N}: FP64 MFMA variant rel L2 vs oracle {err:.2e}")
    assert err < TOL


@pytest.mark.parametrize("N", [1, 2, 4, 5, 6, 7])
def test_maxwell_fp64_matrix_core_variant_vs_oracle(ctx, N):
    mesh = dgc_amd.Mesh.structured(10, 10, -1, 1, -1, 1, jitter=0.1, seed=6)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.MaxwellTM(disc)
    op.waves_per_tile = 0
    assert op.waves_per_tile == 0
    K, Np = disc.K, disc.Np
    u0 = np.random.default_rng(N).standard_normal(3 * K * Np)
    dt = 0.5 * mesh.min_edge / (N + 1) ** 2
    state = upload_fields(ctx, u0, 3, K, Np)
    op.rk4_steps(state, 0.0, dt, 10, op.new_work())
    want, _ = Problem.from_mesh(mesh, N).rk4(0, u0, 0.0, dt, nsteps=10)
    err = rel(download_fields(state, K, Np), want)
    print(f"N={N}: Maxwell FP64 MFMA variant rel L2 vs oracle {err:.2e}")
    assert err < TOL


def test_maxwell_cavity_t1_fp64_matrix_core_variant(ctx, c1):
    fix, mesh, disc, op = c1
    K, Np = disc.K, disc.Np
    n = K * Np
    u0 = np.concatenate([np.zeros(2 * n), np.sin(np.pi * fix["x"]) * np.sin(np.pi * fix["y"])])
    default_wt = op.waves_per_tile
    op.waves_per_tile = 0
    try:
        state = upload_fields(ctx, u0, 3, K, Np)
        t, steps = op.rk4_run(state, 0.0, 1.0, float(fix["dt"]), op.new_work())
        err = rel(download_fields(state, K, Np), fix["final"])
        print(f"maxwell C1 t=1 (FP64 MFMA variant) rel L2 vs reference = {err:.3e}")
        assert steps == int(fix["steps"]) and err < TOL
    finally:
        op.waves_per_tile = default_wt


@pytest.mark.parametrize("N", [1, 3, 7])
def test_maxwell_all_orders_vs_oracle(ctx, N):
    mesh = dgc_amd.Mesh.structured(10, 10, -1, 1, -1, 1, jitter=0.1, seed=6)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.MaxwellTM(disc)
    K, Np = disc.K, disc.Np
    rng = np.random.default_rng(N)
    u0 = rng.standard_normal(3 * K * Np)
    dt = 0.5 * mesh.min_edge / (N + 1) ** 2
    state = upload_fields(ctx, u0, 3, K, Np)
    op.rk4_steps(state, 0.0, dt, 10, op.new_work())
    want, _ = Problem.from_mesh(mesh, N).rk4(0, u0, 0.0, dt, nsteps=10)
    assert rel(download_fields(state, K, Np), want) < TOL


def test_tail_tile_and_odd_sizes(ctx):
    """K not a multiple of the 64-cell tile."""
    mesh = dgc_amd.Mesh.structured(7, 5, -1, 1, -1, 1, jitter=0.1, seed=8)  # K = 70
    for N in (2, 5):
        disc = dgc_amd.Discretization(ctx, mesh, N)
        op = dgc_amd.MaxwellTM(disc)
        K, Np = disc.K, disc.Np
        u = np.random.default_rng(0).standard_normal(3 * K * Np)
        dout = [dgc_amd.DeviceArray(ctx, K * Np) for _ in range(3)]
        op.rhs(upload_fields(ctx, u, 3, K, Np), 0.0, dout)
        assert rel(download_fields(dout, K, Np), Problem.from_mesh(mesh, N).maxwell_rhs(u)) < 1e-13


def test_periodic_mass_conservation_full_config2_size(ctx):
    """Config-2 size K = 131072 at N = 3: mass is conserved to rounding."""
    N = 3
    mesh = dgc_amd.Mesh.structured(256, 256, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.Advection(disc, 1.0, 0.5, dgc_amd.Advection.BC_ZERO)
    K, Np = disc.K, disc.Np
    x, y = mesh.dof_coords(N, node_major=True)
    u0 = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / 0.01)
    M = dgc_amd.ref_operators(N)["M"]
    J = disc.geometry["J"]

    def mass(u):
        return float(np.sum(J * (np.ones(Np) @ M @ u.reshape(Np, K))))

    state = [dgc_amd.DeviceArray.from_host(ctx, u0)]
    dt = 0.5 * mesh.min_edge / (np.hypot(1.0, 0.5) * (N + 1) ** 2)
    op.rk4_steps(state, 0.0, dt, 50, op.new_work())
    u = state[0].download()
    assert np.all(np.isfinite(u))
    assert abs(mass(u) - mass(u0)) < 1e-12 * abs(mass(u0))
    assert np.sum(u**2) <= np.sum(u0**2) * 1.0001  # no growth


@pytest.fixture(scope="module")
def c2_mesh():
    return dgc_amd.Mesh.structured(256, 256, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)


# RMS-relative error of the oracle's own 200-step run against the exact
# translate of the Gaussian on this mesh (measured: 3.66e-4, 5.01e-6,
# 4.33e-8, 3.43e-10, 6.0e-12, 5.5e-12, 5.5e-12 for N = 1..7), with 50 % margin.
C2_EXACT_BOUND = {1: 5.5e-4, 2: 7.5e-6, 3: 6.5e-8, 4: 5.2e-10, 5: 9e-12, 6: 9e-12, 7: 9e-12}


def _gauss_translate(x, y, T, ax, ay):
    out = np.zeros_like(x)
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            out += np.exp(-((x - ax * T - 0.5 + i) ** 2 + (y - ay * T - 0.5 + j) ** 2) / 0.01)
    return out


@pytest.mark.parametrize("N", range(1, 8))
def test_config2_full_size_vs_oracle_and_exact_translation(ctx, c2_mesh, N):
    """The advection workload at its full size (256 x 256 periodic squares,
    K = 131072, a = (1, 0.5), Gaussian, CFL 0.5, 200 RK4 steps) with the
    autotuned tiling: the RHS and the first 5 steps against the oracle
    (FP64 bar), all 200 steps against the oracle for N <= 3, and the final
    state against the exact translate of the Gaussian for every N."""
    mesh = c2_mesh
    print(f"C2 full size N={N}", flush=True)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.Advection(disc, 1.0, 0.5, dgc_amd.Advection.BC_ZERO)
    K, Np = disc.K, disc.Np
    x, y = mesh.dof_coords(N, node_major=False)
    u0 = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / 0.01)
    dt = 0.5 * mesh.min_edge / (np.hypot(1.0, 0.5) * (N + 1) ** 2)
    P = Problem.from_mesh(mesh, N, threaded=True)
    state = upload_fields(ctx, u0, 1, K, Np)
    work = op.new_work()
    op.autotune(state, work, reps=2)
    dout = [dgc_amd.DeviceArray(ctx, K * Np)]
    op.rhs(state, 0.0, dout)
    assert rel(download_fields(dout, K, Np), P.advection_rhs(u0, 0.0, 1.0, 0.5, 1)) < TOL
    op.rk4_steps(state, 0.0, dt, 5, work)
    want5, _ = P.rk4(1, u0, 0.0, dt, nsteps=5, ax=1.0, ay=0.5, bc=1)
    assert rel(download_fields(state, K, Np), want5) < TOL
    op.rk4_steps(state, 5 * dt, dt, 195, work)
    got = download_fields(state, K, Np)
    if N <= 3:
        want, _ = P.rk4(1, u0, 0.0, dt, nsteps=200, ax=1.0, ay=0.5, bc=1)
        assert rel(got, want) < TOL
    err = rel(got, _gauss_translate(x, y, 200 * dt, 1.0, 0.5))
    print(f"N={N} tiling {op.waves_per_tile}: rel L2 vs exact translate after 200 steps {err:.3e}")
    assert err < C2_EXACT_BOUND[N]


def test_time_loops_report_a_nonfinite_state(ctx, c1):
    """A NaN or an infinity in the state (plain data written from the host)
    makes dgc_rk4_steps / dgc_rk4_run return DGC_ERR_NONFINITE; a finite
    state passes the same check."""
    fix, mesh, disc, op = c1
    K, Np = disc.K, disc.Np
    dt = float(fix["dt"])
    work = op.new_work()
    for bad in (np.nan, np.inf):
        u = np.random.default_rng(5).standard_normal(3 * K * Np)
        u[2 * K * Np + 17] = bad  # one Ez value
        state = upload_fields(ctx, u, 3, K, Np)
        with pytest.raises(dgc_amd.DGCError) as ei:
            op.rk4_steps(state, 0.0, dt, 3, work)
        assert ei.value.code == 4 and "NaN" in str(ei.value)
        state = upload_fields(ctx, u, 3, K, Np)
        with pytest.raises(dgc_amd.DGCError) as ei:
            op.rk4_run(state, 0.0, 3.5 * dt, dt, work)
        assert ei.value.code == 4
    state = upload_fields(ctx, np.random.default_rng(5).standard_normal(3 * K * Np), 3, K, Np)
    op.rk4_steps(state, 0.0, dt, 3, work)
    assert np.all(np.isfinite(download_fields(state, K, Np)))


def test_bad_arguments_are_errors(ctx, c1):
    fix, mesh, disc, op = c1
    a = [dgc_amd.DeviceArray(ctx, disc.size) for _ in range(3)]
    with pytest.raises(dgc_amd.DGCError):
        dgc_amd.lib()  # loaded
        from dgc_amd.core import check, _ptrs

        check(dgc_amd.lib().dgc_op_rhs(op._h, _ptrs([x.ptr for x in a]), disc.size + 1, 0.0,
                                       _ptrs([x.ptr for x in a])))
    with pytest.raises(dgc_amd.DGCError):
        op.rk4_run(a, 0.0, 1.0, -1.0, op.new_work())


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 7])
def test_every_tiling_gives_identical_results(ctx, N):
    """The stage kernels' tilings (1, 2, 4, 8 or 16 waves per tile, 32: the
    rows of a tile split over 4 workgroups, and 0: the FP64 matrix-core
    variant) agree to rounding (the summation order and multiply-add
    contraction differ per variant) on a mesh whose last tile is partial."""
    mesh = dgc_amd.Mesh.structured(20, 13, 0, 1, 0, 1, jitter=0.05, seed=4, periodic=True)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    results = []
    for phys in ("adv", "mx"):
        for wt in (1, 2, 4, 0, 8, 16, 32):
            op = (dgc_amd.Advection(disc, 1.0, 0.5, dgc_amd.Advection.BC_ZERO) if phys == "adv"
                  else dgc_amd.MaxwellTM(disc))
            op.waves_per_tile = wt
            F = op.F
            u0 = np.random.default_rng(N).standard_normal(F * disc.size)
            st = [dgc_amd.DeviceArray.from_host(ctx, f) for f in np.split(u0, F)]
            op.rk4_steps(st, 0.0, 1e-4, 3, op.new_work())
            results.append((phys, op.waves_per_tile, np.concatenate([a.download() for a in st])))
    for phys in ("adv", "mx"):
        rs = [r for p, _, r in results if p == phys]
        for r in rs[1:]:
            assert rel(r, rs[0]) < 1e-14


@pytest.mark.parametrize("N", [1, 3, 4])
def test_cache_policies_give_identical_results(ctx, N):
    """The streaming kernels differ from the temporal ones only in the
    nontemporal hint on the accumulator streams: every tiling, both
    operators and the FP32-contraction form give bitwise the same state.
    A small mesh selects the temporal kernels under the automatic policy."""
    mesh = dgc_amd.Mesh.structured(17, 11, 0, 1, 0, 1, jitter=0.05, seed=6, periodic=True)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    cases = [("adv", wt, "fp64") for wt in (1, 2, 4, 0)] + [("mx", wt, "fp64") for wt in (1, 2, 4, 0)]
    cases.append(("adv", 4, "fp32-mfma"))
    for phys, wt, prec in cases:
        out = {}
        for policy in ("auto", "temporal", "streaming"):
            op = (dgc_amd.Advection(disc, 1.0, 0.5, dgc_amd.Advection.BC_ZERO) if phys == "adv"
                  else dgc_amd.MaxwellTM(disc))
            op.waves_per_tile = wt
            op.precision = prec
            op.set_cache_policy(policy)
            assert op.streaming == (policy == "streaming")
            u0 = np.random.default_rng(N).standard_normal(op.F * disc.size)
            st = [dgc_amd.DeviceArray.from_host(ctx, f) for f in np.split(u0, op.F)]
            work = op.new_work()
            op.rk4_steps(st, 0.0, 1e-4, 3, work)  # graph replays
            op.rk4_step(st, 3e-4, 1e-4, work)     # direct launches
            out[policy] = np.concatenate([a.download() for a in st])
        np.testing.assert_array_equal(out["streaming"], out["temporal"], err_msg=f"{phys} wt={wt} {prec}")
        np.testing.assert_array_equal(out["auto"], out["temporal"], err_msg=f"{phys} wt={wt} {prec}")


def test_autotune_leaves_state_untouched(ctx):
    mesh = dgc_amd.Mesh.structured(32, 32, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
    disc = dgc_amd.Discretization(ctx, mesh, 2)
    op = dgc_amd.Advection(disc, 1.0, 0.5, dgc_amd.Advection.BC_ZERO)
    u0 = np.random.default_rng(0).standard_normal(disc.size)
    st = [dgc_amd.DeviceArray.from_host(ctx, u0)]
    res = op.autotune(st, op.new_work(), reps=3)
    assert res["choice"] in (1, 2, 4, 0, 8, 16) and res["ms"] and 0 in res["ms"]
    np.testing.assert_array_equal(st[0].download(), u0)


@pytest.mark.parametrize("n", [1, 2, 7, 1000, (1 << 20) + 3])
def test_axpy_matches_host(ctx, n):
    """dgc_axpy (out = a x + y, the k_axpy replacement) on even, odd and
    16-byte-misaligned vectors: within one rounding of the exact result."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n + 1)
    y = rng.standard_normal(n + 1)
    a = 0.37
    dx, dy, do = (dgc_amd.DeviceArray.from_host(ctx, v) for v in (x, y, np.zeros(n + 1)))
    L = dgc_amd.lib()
    for off in (0, 1):  # off = 1: pointers 8 bytes past a 16-byte boundary
        rc = L.dgc_axpy(ctx.handle, a, dx.ptr + 8 * off, n, dy.ptr + 8 * off, do.ptr + 8 * off)
        assert rc == 0
        got = do.download()[off:off + n]
        want = a * x[off:off + n] + y[off:off + n]
        bound = 2.3e-16 * (np.abs(a * x[off:off + n]) + np.abs(y[off:off + n]))
        assert np.all(np.abs(got - want) <= bound)
