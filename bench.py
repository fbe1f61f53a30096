#!/usr/bin/env python3
"""Benchmark of the FP64 nodal-DG hot path on one MI355X.

Workload (the linear-advection order sweep): u_t + a.grad u = 0, a = (1, 0.5),
periodic unit square, 256 x 256 squares split into K = 131072 triangles with
vertices jittered +-0.05h (seed 1), upwind flux, classical RK4 at CFL 0.5
(dt = 0.5 h_min / (|a| (N+1)^2)), Gaussian pulse initial data, orders
N = 1..7. One bench "step" is one RK4 step of every order in the sweep; the
sweep's whole-job rate is  sum_N (K Np 4 steps) / sum_N (wall time).

Alongside: the roofline of the dominant kernel (HIP events around each stage
kernel; algorithmic bytes per launch), HBM-resident low-order runs (advection
and Maxwell TM at N = 1..3 on 2048 x 2048 squares, K = 8.4M: working sets of
0.8-6 GB, far above the 256 MB Infinity Cache), the reference CPU path
timed on a bounded sample on the host cores, and the IPDG Poisson solvers'
time to FP64 relative residual 1e-10 (block-Jacobi PCG, two-level PCG and
flexible CG with the learned preconditioner, trained on 4096 pairs from the
solver's own iterates) at K = 32768 N = 3 and K = 131072 N = 2 over 256
held-out right-hand sides.

    python bench.py [--steps 200] [--warmup 10] [--gpus 1] [--ipdg-quick]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dg-on-cuda_amd"))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, published
GEO_BYTES_PER_CELL = 4 * 8 + 3 * 4   # metric terms + packed connectivity read per cell per stage
STAGE_PASSES = {1: 3, 2: 5, 3: 4, 4: 5}  # state-vector passes per stage (17 per step)


def stage_bytes(K, Np, F, stage):
    return K * (STAGE_PASSES[stage] * 8 * F * Np + GEO_BYTES_PER_CELL)


def setup_advection(dgc, ctx, mesh, N, a=(1.0, 0.5), setup="host"):
    disc = dgc.Discretization(ctx, mesh, N, setup=setup)
    op = dgc.Advection(disc, a[0], a[1], dgc.Advection.BC_ZERO)
    x, y = mesh.dof_coords(N, node_major=True)
    u0 = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / 0.01)
    state = [dgc.DeviceArray.from_host(ctx, u0)]
    work = op.new_work()
    dt = 0.5 * mesh.min_edge / (np.hypot(*a) * (N + 1) ** 2)
    return disc, op, state, work, dt, u0


def setup_maxwell(dgc, ctx, mesh, N, setup="host"):
    """Maxwell TM, Ez = sin(pi x) sin(pi y), Hx = Hy = 0, dt = 0.5 h_min / (N+1)^2."""
    disc = dgc.Discretization(ctx, mesh, N, setup=setup)
    op = dgc.MaxwellTM(disc)
    x, y = mesh.dof_coords(N, node_major=True)
    ez = np.sin(np.pi * x) * np.sin(np.pi * y)
    state = [dgc.DeviceArray(ctx, disc.size), dgc.DeviceArray(ctx, disc.size), dgc.DeviceArray.from_host(ctx, ez)]
    state[0].zero()
    state[1].zero()
    work = op.new_work()
    dt = 0.5 * mesh.min_edge / (N + 1) ** 2
    return disc, op, state, work, dt, ez


def run_order(dgc, ctx, mesh, N, steps, warmup, prof_steps=5, precision="fp64", physics="advection",
              setup="host"):
    make = setup_maxwell if physics == "maxwell" else setup_advection
    disc, op, state, work, dt, u0 = make(dgc, ctx, mesh, N, setup=setup)
    K, Np, F = disc.K, disc.Np, op.F
    op.precision = precision
    tune = op.autotune(state, work)  # picks the tiling (waves per 64-cell tile)
    if warmup:
        op.rk4_steps(state, 0.0, dt, warmup, work)
    ctx.sync()
    t0 = time.perf_counter()
    ev_ms = op.rk4_steps_timed(state, warmup * dt, dt, steps, work)
    ctx.sync()
    wall = time.perf_counter() - t0
    stage_ms = op.profile_stages(state, 0.0, dt, prof_steps, work)
    u = state[-1].download()
    ok = bool(np.all(np.isfinite(u)))
    dofs = F * K * Np * 4 * steps
    return dict(N=N, Np=Np, K=K, F=F, physics=physics, dt=dt, wall_s=wall, event_ms=ev_ms, dof_stage=dofs,
                value=dofs / wall, stage_ms=stage_ms, finite=ok, tune=tune, streaming=op.streaming, u=u)


MFMA_PEAK_F32_TFLOPS = 157.3  # MI355X dense FP32 matrix, published
MFMA_PEAK_F64_TFLOPS = 78.6   # MI355X dense FP64 matrix, published


def mfma_stats(r, peak=MFMA_PEAK_F32_TFLOPS):
    """Matrix-core rate of a matrix-core stage-2 kernel: issued flops (the
    16x16x4 blocks, padding included) and useful flops (2 Np^2 + Np 3Nfp
    multiply-adds per cell for Dr, Ds and LIFT) per launch over its mean
    duration. The issued count equals SQ_INSTS_VALU_MFMA_MOPS_F32 (F64) x 512
    per launch, and the fraction of peak equals SQ_VALU_MFMA_BUSY_CYCLES per
    SIMD over the kernel's cycles
    (profiles/dg_stage_mfma_kernel_advection_N4-7_K131072_mfma.md)."""
    N, Np, K = r["N"], r["Np"], r["K"]
    NF = 3 * (N + 1)
    MB, KS, KL = (Np + 15) // 16, (Np + 3) // 4, (NF + 3) // 4
    F = r.get("F", 1)  # fields: each gets the Dr, Ds and LIFT products
    issued = 2.0 * 1024 * MB * (2 * KS + KL) * ((K + 15) // 16) * F
    useful = 2.0 * (2 * Np * Np + Np * NF) * K * F
    dur = r["tune"]["ms"][r["tune"]["choice"]] * 1e-3
    return dict(issued_tflops=round(issued / dur / 1e12, 3), useful_tflops=round(useful / dur / 1e12, 3),
                peak_tflops=peak, frac_issued=round(issued / dur / 1e12 / peak, 4))


# HBM traffic per stage-2 launch measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (one run each, scaled x2.000 and x1.000 by the same counters on
# dgc_axpy of known size): profiles/dg_stage_kernel_advection_N7_K131072_pmc.md
# and profiles/dg_stage_kernel_N3_K8388608_hbm.md. Keyed by (physics, N, K,
# waves per tile; 0 = FP64 matrix-core variant); reported as roofline.traffic
# with its source, not measured inside the bench run.
TRAFFIC_SOURCE = "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, profiles/ (not measured in this run)"
MEASURED_TRAFFIC = {("advection", 7, 131072, 16): 196.5e6, ("advection", 7, 131072, 8): 196.8e6,
                    ("advection", 7, 131072, 4): 196.8e6, ("advection", 7, 131072, 2): 197.1e6,
                    ("advection", 7, 131072, 0): 197.2e6,
                    ("advection", 3, 8388608, 2): 3737.7e6, ("maxwell", 3, 8388608, 2): 10754.0e6}


def roofline_for(r):
    """Roofline of the stage-2 kernel (5 state passes) of order r['N']: its mean
    duration over back-to-back launches on the operator's stream (HIP events),
    the figure rocprof's kernel trace reports for it."""
    b = stage_bytes(r["K"], r["Np"], r.get("F", 1), 2)
    dur = r["tune"]["ms"][r["tune"]["choice"]] * 1e-3
    achieved = b / dur / 1e9
    wt = r["tune"]["choice"]
    phys = r.get("physics", "advection")
    traffic = MEASURED_TRAFFIC.get((phys, r["N"], r["K"], wt))
    hint = ", nontemporal streams" if r.get("streaming") else ""
    tiling = "rows split over 4 workgroups of 4 waves" if wt == 32 else f"{wt} waves per tile"
    name = (f"dg_stage_mfma_kernel<double, N={r['N']}, {phys}, stage 2{hint}>" if wt == 0 else
            f"dg_stage_kernel<N={r['N']}, {phys}, stage 2, {tiling}{hint}>")
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                traffic_source=TRAFFIC_SOURCE if traffic else None, kernel=name,
                bytes_per_launch=int(b), avg_launch_us=round(dur * 1e6, 2))


def cpu_baseline(dgc, N=3, steps=40):
    """Reference advection CPU path (serial, as published), bounded sample."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    # The reference mesh class has no periodic wrap: use the same 256x256 jittered
    # lattice on [-1,1]^2 with its inflow boundary (same K, same N, same work per DOF).
    mesh = dgc.Mesh.structured(256, 256, -1, 1, -1, 1, jitter=0.05, seed=1)
    Np = (N + 1) * (N + 2) // 2
    dt = 0.5 * mesh.min_edge / (np.sqrt(2.0) * (N + 1) ** 2)
    if os.path.exists(exe):
        vx, vy, etov = mesh.vertices()
        with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
            np.array([len(vx), etov.size // 3], np.int32).tofile(f)
            vx.tofile(f)
            vy.tofile(f)
            etov.astype(np.int32).tofile(f)
            path = f.name
        try:
            out = subprocess.run([exe, "advection_time", path, str(N), str(steps), repr(dt)],
                                 capture_output=True, text=True, check=True, timeout=300)
        finally:
            os.unlink(path)
        n, secs, _ = out.stdout.split()
        kind = "reference"
    else:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import Problem

        P = Problem.from_mesh(mesh, N)
        x, y = mesh.dof_coords(N, node_major=False)
        u0 = np.sin((x + 1) * np.pi) * np.sin((y + 1) * np.pi)
        t0 = time.perf_counter()
        P.rk4(1, u0, 0.0, dt, nsteps=steps, ax=1.0, ay=1.0, bc=0)
        secs = time.perf_counter() - t0
        kind = "port"
    secs = float(secs)
    value = mesh.K * Np * 4 * steps / secs
    return dict(value=value, unit="DOF-stage updates/s", cores=1, kind=kind,
                sample=f"advection_2d CPU path (a=(1,1), inflow BC), K={mesh.K}, N={N}, {steps} RK4 steps "
                       f"in {secs:.2f} s on 1 host core")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_all_cores(dgc, N=3, steps=100):
    """The restatement of the reference advection path with its cell loops on
    all host cores (OpenMP), same lattice and order as the serial baseline."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Problem

    mesh = dgc.Mesh.structured(256, 256, -1, 1, -1, 1, jitter=0.05, seed=1)
    Np = (N + 1) * (N + 2) // 2
    dt = 0.5 * mesh.min_edge / (np.sqrt(2.0) * (N + 1) ** 2)
    P = Problem.from_mesh(mesh, N, threaded=True)
    x, y = mesh.dof_coords(N, node_major=False)
    u0 = np.sin((x + 1) * np.pi) * np.sin((y + 1) * np.pi)
    P.rk4(1, u0, 0.0, dt, nsteps=2, ax=1.0, ay=1.0, bc=0)
    t0 = time.perf_counter()
    P.rk4(1, u0, 0.0, dt, nsteps=steps, ax=1.0, ay=1.0, bc=0)
    secs = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=mesh.K * Np * 4 * steps / secs, unit="DOF-stage updates/s", cores=cores, kind="port",
                cpu=cpu_model(), nproc=os.cpu_count(),
                sample=f"advection_2d restatement (a=(1,1), inflow BC), K={mesh.K}, N={N}, {steps} RK4 steps in "
                       f"{secs:.2f} s, cell loops on {cores} OpenMP threads")


def stream_probe(dgc, ctx, reps=20):
    """Achievable HBM rate on this box: out = a x + y over 2^27 doubles per
    array (reads 2 GiB, writes 1 GiB; 3 x 1 GiB vectors exceed the 256 MB cache)."""
    n = 1 << 27
    x, y, o = (dgc.DeviceArray(ctx, n) for _ in range(3))
    for a in (x, y, o):
        a.zero()
    call = lambda: dgc.core.check(dgc.lib().dgc_axpy(ctx.handle, 0.5, x.ptr, n, y.ptr, o.ptr))
    for _ in range(3):
        call()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    ctx.sync()
    secs = time.perf_counter() - t0
    gbs = 3 * n * 8 * reps / secs / 1e9
    del x, y, o
    return dict(kernel="axpy2_kernel", bytes_per_launch=3 * n * 8, GBs=round(gbs, 1),
                frac_of_8TBs=round(gbs / HBM_PEAK_GBS, 4))


def maxwell_c1(dgc, ctx, with_cpu=True):
    """Maxwell TM cavity: 16x16 squares on [-1,1]^2 split into K = 512 triangles,
    jitter 0.1h (seed 0), N = 4, dt = 0.5 h_min / (N+1)^2, run to t = 1 with a
    truncated last step, Ez = sin(pi x) sin(pi y); GPU time and, beside it, the
    reference CPU path on the same mesh and steps."""
    N = 4
    mesh = dgc.Mesh.structured(16, 16, -1, 1, -1, 1, jitter=0.1, seed=0)
    disc = dgc.Discretization(ctx, mesh, N)
    op = dgc.MaxwellTM(disc)
    x, y = mesh.dof_coords(N, node_major=True)
    ez = np.sin(np.pi * x) * np.sin(np.pi * y)
    dt = 0.5 * mesh.min_edge / (N + 1) ** 2
    st = [dgc.DeviceArray(ctx, disc.size), dgc.DeviceArray(ctx, disc.size), dgc.DeviceArray.from_host(ctx, ez)]
    st[0].zero()
    st[1].zero()
    work = op.new_work()
    op.rk4_run(st, 0.0, 1.0, dt, work)  # warm-up (graph capture)
    for a, v in zip(st, (None, None, ez)):
        a.zero() if v is None else a.upload(v)
    ctx.sync()
    t0 = time.perf_counter()
    t, steps = op.rk4_run(st, 0.0, 1.0, dt, work)
    ctx.sync()
    secs = time.perf_counter() - t0
    s2 = np.sqrt(2.0)
    exact = np.sin(np.pi * x) * np.sin(np.pi * y) * np.cos(s2 * np.pi)
    err = float(np.mean((st[2].download() - exact) ** 2))
    dofs = 3 * disc.K * disc.Np * 4 * steps
    out = dict(K=disc.K, N=N, steps=steps, t=t, gpu_ms=round(secs * 1e3, 3), gpu_value=dofs / secs,
               ez_mean_square_error_vs_cavity_mode=err, unit="DOF-stage updates/s (3 fields)")
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    if with_cpu and os.path.exists(exe):
        vx, vy, etov = mesh.vertices()
        with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
            np.array([len(vx), etov.size // 3], np.int32).tofile(f)
            vx.tofile(f)
            vy.tofile(f)
            etov.astype(np.int32).tofile(f)
            path = f.name
        try:
            res = subprocess.run([exe, "maxwell_time", path, str(N), str(steps), repr(dt)],
                                 capture_output=True, text=True, check=True, timeout=300)
        finally:
            os.unlink(path)
        csecs = float(res.stdout.split()[1])
        out["cpu_reference"] = dict(value=dofs / csecs, seconds=round(csecs, 3), cores=1, kind="reference")
    return out


def ipdg_bench(dgc, ctx, n_side, N, gx, n_train, epochs, n_eval, n_eval_bj, mixed=False):
    """IPDG Poisson -Lap u = f on [-1,1]^2, u = 0 on the boundary, tau = (N+1)(N+2)/h,
    n_side x n_side squares split into triangles, jitter 0.1h (seed 3). Time to
    FP64 relative residual 1e-10 for block-Jacobi PCG (the first n_eval_bj
    right-hand sides), two-level PCG (gx x gx aggregates) and flexible CG with
    the learned preconditioner on the two-level base (all n_eval), over held-out
    right-hand sides b = M f (f: 2 pi^2 sin sin + 16 Fourier modes with k^-2
    amplitudes, seeds 10000..). The network is first trained on n_train
    (residual, error) pairs produced by the solver itself: 16 recorded iterates
    (every 16 iterations) of two-level PCG solves of n_train / 16 Gaussian
    random-field right-hand sides (k^-2 amplitudes, seeds 100..); the last
    sixteenth of the pairs (whole solves) is held out."""
    mesh = dgc.Mesh.structured(n_side, n_side, -1, 1, -1, 1, jitter=0.1, seed=3)
    disc = dgc.Discretization(ctx, mesh, N)
    op = dgc.IPDG(disc, tau_mode=1)
    t0 = time.perf_counter()
    op.block_jacobi()
    t_bj = time.perf_counter() - t0
    t0 = time.perf_counter()
    nc = op.coarse_space(mesh, gx)
    t_coarse = time.perf_counter() - t0
    net = dgc.LearnedPreconditioner(op, seed=0, base=2)
    t0 = time.perf_counter()
    pairs = net.build_dataset_solver(mesh, n_train // 16, seed0=100, modes=16, snapshots=16)
    t_data = time.perf_counter() - t0
    held = max(1, pairs // 16)
    loss_before = net.eval_loss(pairs - held, held)
    t0 = time.perf_counter()
    hist = net.train(0, pairs - held, epochs=epochs, lr=3e-4, seed=0)
    t_train = time.perf_counter() - t0
    loss_after = net.eval_loss(pairs - held, held)
    print(f"# IPDG K={disc.K} N={N}: coarse {gx}x{gx} ({nc} unknowns) {t_coarse:.2f} s, {pairs} pairs in {t_data:.1f} s, "
          f"training {t_train:.1f} s, held-out energy loss {loss_before:.4f} -> {loss_after:.4f}",
          file=sys.stderr, flush=True)
    xg, yg = mesh.dof_coords(N, node_major=True)
    b = dgc.DeviceArray(ctx, disc.size)
    x = dgc.DeviceArray(ctx, disc.size)
    res = {k: dict(iters=[], ms=[], relres=[]) for k in ("bj_pcg", "two_level_pcg", "learned_fcg")}
    fallbacks = 0
    for i in range(n_eval):
        f, _ = dgc.poisson_rhs(xg, yg, seed=10000 + i)
        op.mass(dgc.DeviceArray.from_host(ctx, f), b)
        runs = (("bj_pcg", lambda: op.pcg(b, x, rtol=1e-10, precond=1)),
                ("two_level_pcg", lambda: op.pcg(b, x, rtol=1e-10, precond=2)),
                ("learned_fcg", lambda: net.fcg(b, x, rtol=1e-10, stall_window=40, stall_factor=0.5)))
        for name, fn in runs:
            if name == "bj_pcg" and i >= n_eval_bj:
                continue
            if i == 0:
                fn()  # graph capture and first-touch outside the timing
            r = fn()
            res[name]["iters"].append(r["iters"])
            res[name]["ms"].append(r["ms"])
            res[name]["relres"].append(r["relres"])
            if name == "learned_fcg" and r["fell_back"] == 1:
                fallbacks += 1
    out = dict(K=disc.K, N=N, n=disc.size, coarse_grid=[gx, gx], coarse_unknowns=nc,
               setup_s=dict(block_jacobi=round(t_bj, 3), coarse=round(t_coarse, 3), dataset=round(t_data, 2),
                            training=round(t_train, 2)),
               training=dict(pairs=pairs - held, held_out=held, solves=n_train // 16, epochs=epochs,
                             rhs_seeds=[100, 100 + n_train // 16 - 1], snapshots_per_solve=16,
                             epoch_loss=[round(v, 5) for v in hist], held_out_loss_untrained=round(loss_before, 5),
                             held_out_loss_trained=round(loss_after, 5)),
               eval_rhs=dict(count=n_eval, seeds=[10000, 10000 + n_eval - 1], bj_pcg_count=min(n_eval, n_eval_bj)),
               fcg_stall_fallbacks=fallbacks)
    for name, v in res.items():
        out[name] = dict(median_iters=int(np.median(v["iters"])), median_ms=round(float(np.median(v["ms"])), 3),
                         mean_ms=round(float(np.mean(v["ms"])), 3), max_relres=float(np.max(v["relres"])))
    out["speedup_learned_vs_bj"] = round(out["bj_pcg"]["median_ms"] / out["learned_fcg"]["median_ms"], 2)
    out["speedup_two_level_vs_bj"] = round(out["bj_pcg"]["median_ms"] / out["two_level_pcg"]["median_ms"], 2)
    out["learned_vs_two_level_ms"] = round(out["learned_fcg"]["median_ms"] / out["two_level_pcg"]["median_ms"], 3)
    out["learned_vs_two_level_iters"] = round(out["learned_fcg"]["median_iters"] / out["two_level_pcg"]["median_iters"], 3)
    out["iteration_bandwidth"] = {
        name: cg_iteration_roofline(disc.K, N, out[name]["median_ms"], out[name]["median_iters"],
                                    nc if name == "two_level_pcg" else 0)
        for name in ("bj_pcg", "two_level_pcg")}
    if mixed:
        f, _ = dgc.poisson_rhs(xg, yg, seed=3)
        op.mass(dgc.DeviceArray.from_host(ctx, f), b)
        op.pcg_mixed(b, x, rtol=1e-10)
        r = op.pcg_mixed(b, x, rtol=1e-10)
        out["mixed_fp32_inner_bj"] = dict(outer=r["outer"], inner_iters=r["inner_iters"], ms=round(r["ms"], 3),
                                         relres=r["relres"])
    for name in ("bj_pcg", "two_level_pcg", "learned_fcg"):
        print(f"#   {name}: median {out[name]['median_iters']} it, {out[name]['median_ms']:.2f} ms, "
              f"max relres {out[name]['max_relres']:.2e}", file=sys.stderr, flush=True)
    del net, op, disc
    return out


def solver_trace(dgc, ctx, n_side=128, N=3):
    """The IPDG solvers of the learned-preconditioner comparison (K = 32768,
    N = 3) with every iteration's kernels launched one by one instead of from
    captured graphs, for kernel traces: set-up, a short training (16 solves x 16
    iterates, one epoch), then one solve each of block-Jacobi PCG, two-level
    PCG and flexible CG with the network."""
    mesh = dgc.Mesh.structured(n_side, n_side, -1, 1, -1, 1, jitter=0.1, seed=3)
    disc = dgc.Discretization(ctx, mesh, N)
    op = dgc.IPDG(disc, tau_mode=1)
    op.set_graphs(False)
    op.block_jacobi()
    op.coarse_space(mesh, 32)
    net = dgc.LearnedPreconditioner(op, seed=0, base=2)
    pairs = net.build_dataset_solver(mesh, 16, seed0=100, modes=16, snapshots=16)
    net.train(0, pairs, epochs=1, lr=3e-4, seed=0)
    xg, yg = mesh.dof_coords(N, node_major=True)
    b = dgc.DeviceArray(ctx, disc.size)
    x = dgc.DeviceArray(ctx, disc.size)
    f, _ = dgc.poisson_rhs(xg, yg, seed=10000)
    op.mass(dgc.DeviceArray.from_host(ctx, f), b)
    out = dict(K=disc.K, N=N, pairs=pairs)
    for name, fn in (("bj_pcg", lambda: op.pcg(b, x, rtol=1e-10, precond=1)),
                     ("two_level_pcg", lambda: op.pcg(b, x, rtol=1e-10, precond=2)),
                     ("learned_fcg", lambda: net.fcg(b, x, rtol=1e-10, stall_window=40, stall_factor=0.5))):
        r = fn()
        out[name] = dict(iters=r["iters"], ms=round(r["ms"], 3), fell_back=r.get("fell_back", 0))
    return out


def ipdg_apply_bytes(K, N):
    """Algorithmic bytes of one SPD apply (pass 1: read u, the 4 metric terms
    and 3 connectivity ints per cell (normals and face_J / J follow from the
    metric terms), write q (2 Np) and the jumps (3 Nfp); pass 2: read q, the
    metric terms, connectivity, penalties (3), jumps and J, write out)."""
    Np, Nfp = (N + 1) * (N + 2) // 2, N + 1
    pass1 = 8 * Np + 4 * 8 + 12 + 16 * Np + 24 * Nfp
    pass2 = 16 * Np + 4 * 8 + 12 + 24 + 24 * Nfp + 8 + 8 * Np
    return K * (pass1 + pass2)


def cg_iteration_roofline(K, N, ms, iters, nc=0):
    """Algorithmic bytes per PCG iteration over its mean time: the apply with
    the direction formed in pass 1 (reads z and the old direction, writes the
    new one) and the dot operand in pass 2, the update (reads r, x, p, q and
    the packed block inverses, writes x, r, z); two-level adds restriction
    (r, 6 FP32 coefficients), the coarse solve (FP32 inverse, nc x nc) and the
    prolongation (reads z, r, writes z)."""
    Np = (N + 1) * (N + 2) // 2
    b = ipdg_apply_bytes(K, N) + K * (8 * Np + 8 * Np + 8 * Np)  # z + new direction + dot operand
    b += K * (32 * Np + 8 * Np * (Np + 1) // 2 + 24 * Np)
    if nc:
        b += K * (8 * Np + 24) + 4 * nc * nc + K * 24 * Np
    us = ms * 1e3 / max(1, iters)
    return dict(bytes_per_iteration=int(b), us_per_iteration=round(us, 2), GBs=round(b / us / 1e3, 1),
                frac_of_8TBs=round(b / us / 1e3 / HBM_PEAK_GBS, 4))


def ipdg_apply_bench(dgc, ctx, reps=30):
    """The SPD IPDG apply (both passes) at N = 3 on 256 x 256 (K = 131072,
    cache-resident) and 1024 x 1024 squares (K = 2,097,152: vectors of 168 MB,
    HBM-resident), tau = (N+1)(N+2)/h, random input; wall time of `reps`
    back-to-back applies on the solver's stream."""
    out = []
    for n_side in (256, 1024):
        mesh = dgc.Mesh.structured(n_side, n_side, -1, 1, -1, 1, jitter=0.1, seed=3)
        disc = dgc.Discretization(ctx, mesh, 3)
        op = dgc.IPDG(disc, tau_mode=1)
        u = dgc.DeviceArray.from_host(ctx, np.random.default_rng(0).standard_normal(disc.size))
        w = dgc.DeviceArray(ctx, disc.size)
        for _ in range(3):
            op.apply_spd(u, w)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            op.apply_spd(u, w)
        ctx.sync()
        us = (time.perf_counter() - t0) / reps * 1e6
        b = ipdg_apply_bytes(disc.K, 3)
        out.append(dict(K=disc.K, N=3, us_per_apply=round(us, 2), bytes_per_apply=int(b), GBs=round(b / us / 1e3, 1),
                        frac_of_8TBs=round(b / us / 1e3 / HBM_PEAK_GBS, 4)))
        print(f"# IPDG SPD apply K={disc.K} N=3: {us:.1f} us, {out[-1]['GBs']} GB/s", file=sys.stderr, flush=True)
        del op, disc, u, w
    return out


def ensemble_bench(dgc, ctx, n_inst=64, steps=500, seed=2):
    """64 independent FP64 Maxwell TM cavities on [-1,1]^2: K log-uniform in
    [2e3, 2e5] (structured squares split into triangles, jitter 0.1h), N uniform
    in {2..6}, Ez = sin(m pi x) sin(n pi y) with m, n uniform in {1..4}, all
    drawn from seed 2; dt = 0.5 h_min / (N+1)^2 per instance. The ensemble
    (one launch per order and RK4 stage for all instances, the step replayed
    from a HIP graph, tilings autotuned per order) against the same steps run
    instance after instance with the per-order kernels."""
    rng = np.random.default_rng(seed)
    ops, states, works, dts, dofs, orders, Ks = [], [], [], [], 0, [], []
    step_bytes = 0  # algorithmic bytes of one RK4 step of every instance
    t0 = time.perf_counter()
    for i in range(n_inst):
        k_target = float(np.exp(rng.uniform(np.log(2e3), np.log(2e5))))
        N = int(rng.integers(2, 7))
        m, n = (int(v) for v in rng.integers(1, 5, size=2))
        side = max(2, int(round(np.sqrt(k_target / 2))))
        mesh = dgc.Mesh.structured(side, side, -1, 1, -1, 1, jitter=0.1, seed=1000 + i)
        disc = dgc.Discretization(ctx, mesh, N)
        op = dgc.MaxwellTM(disc)
        x, y = mesh.dof_coords(N, node_major=True)
        ez = np.sin(m * np.pi * x) * np.sin(n * np.pi * y)
        st = [dgc.DeviceArray(ctx, disc.size), dgc.DeviceArray(ctx, disc.size), dgc.DeviceArray.from_host(ctx, ez)]
        st[0].zero()
        st[1].zero()
        ops.append(op)
        states.append(st)
        works.append(op.new_work())
        dts.append(0.5 * mesh.min_edge / (N + 1) ** 2)
        dofs += 3 * disc.K * disc.Np
        step_bytes += sum(stage_bytes(disc.K, disc.Np, 3, st_) for st_ in (1, 2, 3, 4))
        orders.append(N)
        Ks.append(disc.K)
    t_setup = time.perf_counter() - t0
    ens = dgc.Ensemble(ctx, ops, states, works, dts)
    ens.autotune(reps=10)
    info = ens.info()
    ens.steps(5)
    ms_e = ens.steps(steps)
    ens.steps_sequential(2)
    ms_s = ens.steps_sequential(steps)
    finite = all(bool(np.all(np.isfinite(s[2].download()))) for s in states)
    v_e = dofs * 4 * steps / (ms_e * 1e-3)
    v_s = dofs * 4 * steps / (ms_s * 1e-3)
    wt_by_order = {}
    for N, w in zip(orders, info["waves_per_tile"]):
        wt_by_order[N] = w
    out = dict(instances=n_inst, total_dofs=dofs, steps=steps, K_min=min(Ks), K_max=max(Ks),
               orders={N: orders.count(N) for N in sorted(set(orders))},
               kernel_by_order={N: ("FP64 MFMA, 64 cells per workgroup" if w == 0 else
                                    f"VALU, {64 * max(1, 4 // w)} cells per workgroup of {max(4, w)} waves")
                                for N, w in sorted(wt_by_order.items())},
               work_items=info["num_items"], grid=info["grid"], lds_bytes=info["lds_bytes"],
               nontemporal_streams=ens.streaming,
               step_bytes=step_bytes,
               ensemble=dict(value=v_e, ms_per_step=ms_e / steps,
                             GBs=round(step_bytes / (ms_e / steps * 1e-3) / 1e9, 1),
                             frac_of_8TBs=round(step_bytes / (ms_e / steps * 1e-3) / 8e12, 4)),
               sequential=dict(value=v_s, ms_per_step=ms_s / steps),
               speedup=round(v_e / v_s, 3), finite=finite, setup_s=round(t_setup, 1),
               unit="DOF-stage updates/s (3 fields)")
    print(f"# ensemble of {n_inst} Maxwell instances ({dofs / 1e6:.1f} M DOFs): {v_e / 1e9:.2f} G DOF-stage/s "
          f"vs sequential {v_s / 1e9:.2f} ({out['speedup']}x)", file=sys.stderr, flush=True)
    del ens
    return out


def graddiv_bench(dgc, ctx, N=3, reps=30):
    """grad_2d and div_2d with central fluxes (dgc_grad2d_apply / dgc_div2d_apply)
    on 1024 x 1024 squares (K = 2,097,152, jitter 0.05h, seed 1) at order N:
    mean time per launch over `reps` back-to-back launches on the context's
    stream (wall clock around them, launch overhead included). Algorithmic
    HBM bytes per cell: grad reads u and writes two components, div reads two
    and writes one, both 24 Np + 44 (8 B per node value, four metric terms,
    three connectivity words)."""
    mesh = dgc.Mesh.structured(1024, 1024, 0, 1, 0, 1, jitter=0.05, seed=1)
    disc = dgc.Discretization(ctx, mesh, N, setup="device")
    op = dgc.GradDiv(disc)
    n = disc.size
    x, y = mesh.dof_coords(N, node_major=True)
    u0 = dgc.DeviceArray.from_host(ctx, np.sin(np.pi * x) * np.cos(np.pi * y))
    u1 = dgc.DeviceArray.from_host(ctx, np.cos(np.pi * x) * np.sin(np.pi * y))
    o0, o1 = dgc.DeviceArray(ctx, n), dgc.DeviceArray(ctx, n)
    out = {}
    bytes_per = (24 * disc.Np + 44) * disc.K
    for name, call in (("grad", lambda: op.grad(u0, o0, o1)), ("div", lambda: op.div(u0, u1, o0))):
        for _ in range(3):
            call()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        ctx.sync()
        dur = (time.perf_counter() - t0) / reps
        gbs = bytes_per / dur / 1e9
        out[name] = dict(us=round(dur * 1e6, 1), GBs=round(gbs, 1), frac_of_peak=round(gbs / HBM_PEAK_GBS, 4))
    out.update(K=disc.K, N=N, bytes_per_launch=bytes_per)
    print(f"# grad/div K={disc.K} N={N}: grad {out['grad']['us']} us ({out['grad']['GBs']} GB/s), "
          f"div {out['div']['us']} us ({out['div']['GBs']} GB/s)", file=sys.stderr, flush=True)
    return out


def stokes_bench(dgc, ctx):
    """The Stokes example's run on the GPU (examples/stokes_2d/main.cpp:65-196):
    the reference's fine mesh (tests/golden/square_domain_fine.msh, K = 1064),
    dt = 1e-3 to T = 0.1 from u^{n-1} = u^n = exact, at N = 1 (the example's
    default) and N = 2; wall time of the 100 steps, mean-square error against the
    exact solution as the example prints it, iterations of the last step. Then
    the step time on 128 x 128 squares (K = 32,768) at N = 2 over 5 steps, with
    block-Jacobi and with two-level (32 x 32 boxes) pressure solves."""
    from dgc_amd.stokes import exact_velocity

    out = {}
    fine = dgc.Mesh.read_gmsh(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                            "square_domain_fine.msh"))
    big = dgc.Mesh.structured(128, 128, -1, 1, -1, 1, jitter=0.05, seed=1)
    for name, mesh, N, steps, coarse in (("fine_N1", fine, 1, 100, 0), ("fine_N2", fine, 2, 100, 0),
                                         ("K32768_N2", big, 2, 5, 0), ("K32768_N2_two_level", big, 2, 5, 32)):
        disc = dgc.Discretization(ctx, mesh, N)
        st = dgc.Stokes(disc)
        if coarse:
            st.coarse_space(mesh, coarse)
        x, y = mesh.dof_coords(N, node_major=True)
        ux, uy = exact_velocity(x, y, 0.0)
        up = (dgc.DeviceArray.from_host(ctx, ux), dgc.DeviceArray.from_host(ctx, uy))
        u = (dgc.DeviceArray.from_host(ctx, ux), dgc.DeviceArray.from_host(ctx, uy))
        nxt = (dgc.DeviceArray(ctx, disc.size), dgc.DeviceArray(ctx, disc.size))
        tp = t = 0.0
        dt = 1e-3
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            its = st.advance(up, tp, u, t, dt, nxt, 1e-11)
            up, u, nxt = u, nxt, up
            tp, t = t, t + dt
        ctx.sync()
        wall = time.perf_counter() - t0
        ex, ey = exact_velocity(x, y, t)
        err = float((np.sum((u[0].download() - ex) ** 2) + np.sum((u[1].download() - ey) ** 2)) / len(x))
        out[name] = dict(K=disc.K, N=N, steps=steps, T=round(t, 6), wall_s=round(wall, 4),
                         ms_per_step=round(wall * 1e3 / steps, 3), error_norm=err, last_step_iterations=its)
        print(f"# stokes {name}: {out[name]['ms_per_step']} ms/step, error norm {err:.3e}, iterations {its}",
              file=sys.stderr, flush=True)
        del st, disc
    return out


def assemble_bench(dgc, ctx, N=3):
    """The IPDG Laplacian as a CSR matrix (dgc_ipdg_assemble) on 128 x 128 squares
    (K = 32,768) at N = 3: wall time of the assembly (host lists + GPU probing +
    download of the values) against the time the reference's unit-vector
    assembly (examples/stokes_2d/stokes_2d.h:388-422) would take with this
    library's own apply, i.e. n applications."""
    mesh = dgc.Mesh.structured(128, 128, -1, 1, -1, 1, jitter=0.05, seed=1)
    disc = dgc.Discretization(ctx, mesh, N)
    ip = dgc.IPDG(disc, tau_mode=1)
    ip.assemble()  # warm-up
    t0 = time.perf_counter()
    rp, ci, va = ip.assemble()
    wall = time.perf_counter() - t0
    u = dgc.DeviceArray.from_host(ctx, np.random.default_rng(0).standard_normal(disc.size))
    for _ in range(5):
        ip.apply(u)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(50):
        ip.apply(u)
    ctx.sync()
    apply_s = (time.perf_counter() - t0) / 50
    out = dict(K=disc.K, N=N, n=disc.size, nnz=int(len(va)), wall_s=round(wall, 4),
               apply_us=round(apply_s * 1e6, 2), unit_vector_estimate_s=round(apply_s * disc.size, 2))
    print(f"# assemble: {out}", file=sys.stderr, flush=True)
    return out


def adv1d_bench(dgc, ctx):
    """The 1D example's run (examples/advection_1d/main.cpp:83-116): 8192 cells,
    order 6, 10000 RK4 steps of dt = 0.25 / N^2 h / a from sin(x); wall time,
    mean-square error against sin(x - 2 pi t) as the example prints it, and the
    stage kernels' bandwidth: per cell and stage 8 (N+1) x (reads of the stage
    input, u, and the RK streams; writes of the outputs): stages 1-3 move 4
    vectors, stage 4 five."""
    K, N, steps = 8192, 6, 10000
    D, L, x = dgc.Advection1D.operators(K, N)
    op = dgc.Advection1D(ctx, K, N)
    u = dgc.DeviceArray.from_host(ctx, np.sin(x))
    dt = 0.25 / N / N * (2 * np.pi / K) / (2 * np.pi)
    op.rk4_steps(u, 0.0, dt, 10)  # warm-up (not part of the example's run)
    u = dgc.DeviceArray.from_host(ctx, np.sin(x))
    ctx.sync()
    t0 = time.perf_counter()
    op.rk4_steps(u, 0.0, dt, steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    err = float(np.mean((u.download() - np.sin(x - 2 * np.pi * steps * dt)) ** 2))
    out = dict(K=K, N=N, steps=steps, wall_ms=round(wall * 1e3, 2), us_per_step=round(wall * 1e6 / steps, 2),
               error_norm=err, dof_stage_per_s=K * (N + 1) * 4 * steps / wall)
    print(f"# advection 1D K={K} N={N}: {steps} steps in {out['wall_ms']} ms ({out['us_per_step']} us/step), "
          f"error norm {err:.3e}", file=sys.stderr, flush=True)
    return out


def maxwell_fp32_form(dgc, ctx, steps=50, warmup=5):
    """Maxwell TM on the sweep's 256 x 256 periodic lattice (K = 131072),
    N = 4..7, Ez = sin(pi x) sin(pi y): the FP64 path (autotuned tiling)
    against the FP32-contraction form, rate and relative L2 difference of Ez
    after the same steps."""
    mesh = dgc.Mesh.structured(256, 256, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
    out = []
    for N in (4, 5, 6, 7):
        r = run_order(dgc, ctx, mesh, N, steps, warmup, prof_steps=2, physics="maxwell")
        q = run_order(dgc, ctx, mesh, N, steps, warmup, prof_steps=2, physics="maxwell", precision="fp32-mfma")
        err = float(np.linalg.norm(q["u"] - r["u"]) / np.linalg.norm(r["u"]))
        out.append(dict(N=N, fp64=r["value"], fp64_tiling=r["tune"]["choice"], fp32_mfma=q["value"],
                        speedup=round(q["value"] / r["value"], 3), rel_l2_vs_fp64=err,
                        fp32_mfma_matrix=mfma_stats(q)))
        print(f"# Maxwell N={N}: fp64 {r['value'] / 1e9:.1f} G, fp32-mfma {q['value'] / 1e9:.1f} G "
              f"({out[-1]['speedup']}x), rel L2 {err:.1e}", file=sys.stderr, flush=True)
        del r, q
    return out


def setup_timing(dgc, ctx, mesh, N):
    """Wall time of building the device discretization of `mesh` at order N:
    host set-up (geometry and connectivity in C++, then upload) against
    device set-up (upload of the cell corners, geometry, half-edge sort and
    packing on the GPU, host copies back). Both give the same arrays bit for
    bit (tests/test_gpu_setup.py)."""
    ctx.sync()
    t0 = time.perf_counter()
    d = dgc.Discretization(ctx, mesh, N, setup="host")
    ctx.sync()
    th = time.perf_counter() - t0
    del d
    t0 = time.perf_counter()
    d = dgc.Discretization(ctx, mesh, N, setup="device")
    ctx.sync()
    td = time.perf_counter() - t0
    del d
    return dict(K=mesh.K, N=N, host_s=round(th, 3), device_s=round(td, 3), speedup=round(th / td, 1))


def hbm_resident(dgc, ctx, steps, warmup):
    """Low-order FP64 RK4 with every vector in HBM: 2048 x 2048 periodic
    squares (K = 8,388,608, jitter 0.05h, seed 1), advection (a = (1, 0.5),
    Gaussian pulse) and Maxwell TM (Ez = sin(pi x) sin(pi y)) at N = 1, 2, 3.
    The smallest working set (advection N = 1: 5 vectors of 201 MB plus
    369 MB of geometry and connectivity) is 5.4x the 256 MB Infinity Cache,
    so every stage streams from HBM; the automatic cache policy runs the
    nontemporal-stream kernels."""
    hm = dgc.Mesh.structured(2048, 2048, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
    out = []
    for physics in ("advection", "maxwell"):
        for N in (1, 2, 3):
            r = run_order(dgc, ctx, hm, N, steps, warmup, prof_steps=3, physics=physics, setup="device")
            roof = roofline_for(r)
            out.append(dict(physics=physics, N=N, K=r["K"], value=r["value"], finite=r["finite"],
                            stage_us=[round(v * 1e3, 1) for v in r["stage_ms"]],
                            step_GBs=round(sum(stage_bytes(r["K"], r["Np"], r["F"], s) for s in (1, 2, 3, 4))
                                           / (sum(r["stage_ms"]) * 1e-3) / 1e9, 1),
                            roofline=roof))
            print(f"# HBM-resident {physics} K={r['K']} N={N}: {r['value'] / 1e9:.2f} G DOF-stage/s, "
                  f"stage-2 {roof['achieved']} GB/s ({roof['frac']:.1%}), whole step {out[-1]['step_GBs']} GB/s",
                  file=sys.stderr, flush=True)
            del r
    st = setup_timing(dgc, ctx, hm, 3)
    print(f"# set-up K={st['K']} N=3: host {st['host_s']} s, device {st['device_s']} s", file=sys.stderr, flush=True)
    return out, st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--orders", default="1,2,3,4,5,6,7")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-hbm", action="store_true")
    ap.add_argument("--hbm-only", action="store_true",
                    help="run only the HBM-resident N = 3 case (for counter collection)")
    ap.add_argument("--hbm-physics", default="advection", choices=("advection", "maxwell"))
    ap.add_argument("--ipdg-apply-only", action="store_true",
                    help="run only the IPDG SPD apply measurement (for kernel traces and counters)")
    ap.add_argument("--calib", action="store_true",
                    help="also run dgc_axpy over 2^27 doubles: a kernel of known traffic "
                         "(reads 2 GiB, writes 1 GiB) to scale HBM byte counters")
    ap.add_argument("--no-ipdg", action="store_true", help="skip the IPDG solver comparison")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 matrix-core form of the sweep")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched Maxwell ensemble")
    ap.add_argument("--no-graddiv", action="store_true",
                    help="skip the gradient / divergence, Stokes and 1D advection timings")
    ap.add_argument("--no-stokes", action="store_true",
                    help="with --widened-only: skip the Stokes run (its graph-replayed pressure solves cannot be "
                         "kernel-traced)")
    ap.add_argument("--widened-only", action="store_true",
                    help="run only the gradient / divergence, Stokes and 1D advection timings (for kernel traces)")
    ap.add_argument("--solver-trace", action="store_true",
                    help="run only the IPDG solvers (K = 32768, N = 3) with uncaptured launches (for kernel traces)")
    ap.add_argument("--maxwell-order", type=int, default=0,
                    help="run only Maxwell TM at this order on the 256 x 256 mesh (for counter collection)")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--ipdg-quick", action="store_true",
                    help="IPDG comparison on a small sample (512 training pairs, 8 held-out right-hand sides) "
                         "instead of the stated 4096 / 256")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described, not built)"}))
        sys.exit(2)

    import dgc_amd as dgc

    ctx = dgc.Context(0)
    if args.calib:
        n = 1 << 27
        x, y, o = (dgc.DeviceArray(ctx, n) for _ in range(3))
        for a in (x, y, o):
            a.zero()
        for _ in range(3):
            dgc.core.check(dgc.lib().dgc_axpy(ctx.handle, 0.5, x.ptr, n, y.ptr, o.ptr))
        ctx.sync()
        print("# calibration: 3 x dgc_axpy over 2^27 doubles done", file=sys.stderr, flush=True)
        del x, y, o
    if args.widened_only:
        out = dict(workload="gradient / divergence, Stokes step, 1D advection, IPDG assembly",
                   graddiv=graddiv_bench(dgc, ctx), advection_1d=adv1d_bench(dgc, ctx),
                   assemble=assemble_bench(dgc, ctx))
        if not args.no_stokes:
            out["stokes"] = stokes_bench(dgc, ctx)
        print(json.dumps(out))
        return
    if args.solver_trace:
        print(json.dumps(dict(workload="ipdg solvers K=32768 N=3, uncaptured", **solver_trace(dgc, ctx))))
        return
    if args.maxwell_order:
        mesh = dgc.Mesh.structured(256, 256, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
        r = run_order(dgc, ctx, mesh, args.maxwell_order, args.steps, args.warmup, physics="maxwell")
        print(json.dumps(dict(workload=f"maxwell fp64 rk4 K={r['K']} N={args.maxwell_order}", value=r["value"],
                              waves_per_tile=r["tune"]["choice"], roofline=roofline_for(r),
                              stage_us=[round(v * 1e3, 2) for v in r["stage_ms"]])))
        return
    if args.ipdg_apply_only:
        print(json.dumps(dict(workload="ipdg spd apply N=3", ipdg_apply=ipdg_apply_bench(dgc, ctx))))
        return
    if args.hbm_only:
        hm = dgc.Mesh.structured(2048, 2048, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
        r = run_order(dgc, ctx, hm, 3, args.steps, args.warmup, physics=args.hbm_physics, setup="device")
        print(json.dumps(dict(workload=f"{args.hbm_physics} fp64 rk4 K={r['K']} N=3 (HBM-resident)",
                              value=r["value"], roofline=roofline_for(r),
                              stage_us=[round(v * 1e3, 2) for v in r["stage_ms"]])))
        return
    mesh = dgc.Mesh.structured(256, 256, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
    orders = [int(o) for o in args.orders.split(",")]
    per = []
    fp32 = {}
    for N in orders:
        r = run_order(dgc, ctx, mesh, N, args.steps, args.warmup)
        per.append(r)
        print(f"# N={N} Np={r['Np']}: {r['value'] / 1e9:.2f} G DOF-stage/s, "
              f"{r['wall_s'] * 1e3 / args.steps:.3f} ms/step, stage us "
              f"{[round(x * 1e3, 1) for x in r['stage_ms']]}", file=sys.stderr, flush=True)
        if not args.no_fp32:
            # FP32 matrix-core form: same steps from the same initial state
            q = run_order(dgc, ctx, mesh, N, args.steps, args.warmup, precision="fp32-mfma")
            err = float(np.linalg.norm(q["u"] - r["u"]) / np.linalg.norm(r["u"]))
            fp32[N] = dict(value=q["value"], ms_per_step=q["wall_s"] * 1e3 / args.steps,
                           stage_us=[round(x * 1e3, 2) for x in q["stage_ms"]],
                           stage2_GBs=roofline_for(q)["achieved"], rel_l2_vs_fp64=err,
                           speedup_vs_fp64=round(q["value"] / r["value"], 3), mfma=mfma_stats(q))
            print(f"#   fp32-mfma: {q['value'] / 1e9:.2f} G DOF-stage/s ({fp32[N]['speedup_vs_fp64']}x), "
                  f"rel L2 vs fp64 after {args.warmup + args.steps + 5} steps {err:.2e}, "
                  f"MFMA {fp32[N]['mfma']['issued_tflops']} TFLOP/s issued", file=sys.stderr, flush=True)
            del q
        r.pop("u")
    total_dofs = sum(r["dof_stage"] for r in per)
    total_s = sum(r["wall_s"] for r in per)
    value = total_dofs / total_s
    dom = max(per, key=lambda r: r["wall_s"])
    roof = roofline_for(dom)

    extra = {}
    if not args.no_hbm:
        extra["hbm_resident"], extra["setup"] = hbm_resident(dgc, ctx, max(10, args.steps // 20), 2)
    if not args.no_fp32:
        extra["maxwell_fp32_mfma"] = maxwell_fp32_form(dgc, ctx)
    if not args.no_graddiv:
        extra["graddiv"] = graddiv_bench(dgc, ctx)
        extra["stokes"] = stokes_bench(dgc, ctx)
        extra["advection_1d"] = adv1d_bench(dgc, ctx)
        extra["assemble"] = assemble_bench(dgc, ctx)

    if not args.no_ipdg:
        quick = args.ipdg_quick
        extra["ipdg_apply"] = ipdg_apply_bench(dgc, ctx)
        # the learned-preconditioner configuration: 4096 training pairs, 256
        # held-out right-hand sides (block-Jacobi PCG on the first 32 of them)
        extra["ipdg"] = [
            ipdg_bench(dgc, ctx, 128, 3, 32, 512 if quick else 4096, 1 if quick else 3, 8 if quick else 256,
                       8 if quick else 32, mixed=True),
            ipdg_bench(dgc, ctx, 256, 2, 32, 512 if quick else 4096, 1 if quick else 3, 8 if quick else 256,
                       8 if quick else 32),
        ]

    if not args.no_ensemble:
        extra["ensemble"] = ensemble_bench(dgc, ctx, steps=args.ensemble_steps)

    extra["hbm_stream_probe"] = stream_probe(dgc, ctx)
    print(f"# STREAM-like probe: {extra['hbm_stream_probe']['GBs']} GB/s", file=sys.stderr, flush=True)
    extra["maxwell_c1"] = maxwell_c1(dgc, ctx, with_cpu=not args.no_cpu)
    print(f"# Maxwell C1: {extra['maxwell_c1']['gpu_ms']} ms on the GPU for {extra['maxwell_c1']['steps']} steps",
          file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(dgc)
        print(f"# CPU baseline: {cpu['value'] / 1e6:.2f} M DOF-stage/s ({cpu['kind']})", file=sys.stderr)
        extra["cpu_baseline_all_cores"] = cpu_baseline_all_cores(dgc)
        print(f"# CPU baseline, all cores: {extra['cpu_baseline_all_cores']['value'] / 1e6:.2f} M DOF-stage/s "
              f"({extra['cpu_baseline_all_cores']['cores']} threads)", file=sys.stderr, flush=True)

    line = {
        "metric": "DOF-stage updates/s",
        "value": value,
        "unit": "DOF-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_s * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "advection2d_fp64_rk4_order_sweep", "K": mesh.K, "orders": orders,
                   "a": [1.0, 0.5], "mesh": "256x256 periodic, jitter 0.05h seed 1", "cfl": 0.5,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "per_order": [dict(N=r["N"], value=r["value"], ms_per_step=r["wall_s"] * 1e3 / args.steps,
                           stage_us=[round(x * 1e3, 2) for x in r["stage_ms"]],
                           stage2_GBs=roofline_for(r)["achieved"], finite=r["finite"],
                           waves_per_tile=r["tune"]["choice"],
                           fp64_mfma=(mfma_stats(r, MFMA_PEAK_F64_TFLOPS) if r["tune"]["choice"] == 0 else None),
                           autotune_stage2_us={k: round(v * 1e3, 2) for k, v in r["tune"]["ms"].items()},
                           fp32_mfma=fp32.get(r["N"]))
                      for r in per],
        **extra,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
