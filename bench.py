#!/usr/bin/env python3
"""Benchmark of the FP64 nodal-DG hot path on one MI355X.

Workload (the linear-advection order sweep): u_t + a.grad u = 0, a = (1, 0.5),
periodic unit square, 256 x 256 squares split into K = 131072 triangles with
vertices jittered +-0.05h (seed 1), upwind flux, classical RK4 at CFL 0.5
(dt = 0.5 h_min / (|a| (N+1)^2)), Gaussian pulse initial data, orders
N = 1..7. One bench "step" is one RK4 step of every order in the sweep; the
sweep's whole-job rate is  sum_N (K Np 4 steps) / sum_N (wall time).

Alongside: the roofline of the dominant kernel (HIP events around each stage
kernel; algorithmic bytes per launch), an HBM-resident variant (K = 2.1M at
N = 3, larger than the 256 MB Infinity Cache), and the reference CPU path
timed on a bounded sample on the host cores.

    python bench.py [--steps 200] [--warmup 10] [--gpus 1]
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


def setup_advection(dgc, ctx, mesh, N, a=(1.0, 0.5)):
    disc = dgc.Discretization(ctx, mesh, N)
    op = dgc.Advection(disc, a[0], a[1], dgc.Advection.BC_ZERO)
    x, y = mesh.dof_coords(N, node_major=True)
    u0 = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / 0.01)
    state = [dgc.DeviceArray.from_host(ctx, u0)]
    work = op.new_work()
    dt = 0.5 * mesh.min_edge / (np.hypot(*a) * (N + 1) ** 2)
    return disc, op, state, work, dt, u0


def run_order(dgc, ctx, mesh, N, steps, warmup, prof_steps=5):
    disc, op, state, work, dt, u0 = setup_advection(dgc, ctx, mesh, N)
    K, Np = disc.K, disc.Np
    tune = op.autotune(state, work)  # picks the tiling (waves per 64-cell tile)
    if warmup:
        op.rk4_steps(state, 0.0, dt, warmup, work)
    ctx.sync()
    t0 = time.perf_counter()
    ev_ms = op.rk4_steps_timed(state, warmup * dt, dt, steps, work)
    ctx.sync()
    wall = time.perf_counter() - t0
    stage_ms = op.profile_stages(state, 0.0, dt, prof_steps, work)
    u = state[0].download()
    ok = bool(np.all(np.isfinite(u)))
    dofs = K * Np * 4 * steps
    return dict(N=N, Np=Np, K=K, dt=dt, wall_s=wall, event_ms=ev_ms, dof_stage=dofs,
                value=dofs / wall, stage_ms=stage_ms, finite=ok, tune=tune)


def roofline_for(r, F=1):
    """Roofline of the stage-2 kernel (5 state passes) of order r['N']: its mean
    duration over back-to-back launches on the operator's stream (HIP events),
    the figure rocprof's kernel trace reports for it."""
    b = stage_bytes(r["K"], r["Np"], F, 2)
    dur = r["tune"]["ms"][r["tune"]["choice"]] * 1e-3
    achieved = b / dur / 1e9
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                kernel=f"dg_stage_kernel<N={r['N']}, advection, stage 2>",
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--orders", default="1,2,3,4,5,6,7")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-hbm", action="store_true")
    ap.add_argument("--hbm-only", action="store_true",
                    help="run only the HBM-resident case (for counter collection)")
    ap.add_argument("--calib", action="store_true",
                    help="also run dgc_axpy over 2^27 doubles: a kernel of known traffic "
                         "(reads 2 GiB, writes 1 GiB) to scale HBM byte counters")
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
    if args.hbm_only:
        hm = dgc.Mesh.structured(1024, 1024, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
        r = run_order(dgc, ctx, hm, 3, args.steps, args.warmup)
        print(json.dumps(dict(workload="advection2d_fp64_rk4 K=2097152 N=3 (HBM-resident)",
                              value=r["value"], roofline=roofline_for(r),
                              stage_us=[round(v * 1e3, 2) for v in r["stage_ms"]])))
        return
    mesh = dgc.Mesh.structured(256, 256, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
    orders = [int(o) for o in args.orders.split(",")]
    per = []
    for N in orders:
        r = run_order(dgc, ctx, mesh, N, args.steps, args.warmup)
        per.append(r)
        print(f"# N={N} Np={r['Np']}: {r['value'] / 1e9:.2f} G DOF-stage/s, "
              f"{r['wall_s'] * 1e3 / args.steps:.3f} ms/step, stage us "
              f"{[round(x * 1e3, 1) for x in r['stage_ms']]}", file=sys.stderr, flush=True)
    total_dofs = sum(r["dof_stage"] for r in per)
    total_s = sum(r["wall_s"] for r in per)
    value = total_dofs / total_s
    dom = max(per, key=lambda r: r["wall_s"])
    roof = roofline_for(dom)

    extra = {}
    if not args.no_hbm:
        hm = dgc.Mesh.structured(1024, 1024, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
        r = run_order(dgc, ctx, hm, 3, max(10, args.steps // 10), 2)
        extra["hbm_resident"] = dict(K=r["K"], N=3, value=r["value"], roofline=roofline_for(r),
                                     note="2.1M cells x Np=10: 168 MB per state vector, above the "
                                          "256 MB Infinity Cache with the 5 RK buffers")
        print(f"# HBM-resident K={r['K']} N=3: {r['value'] / 1e9:.2f} G DOF-stage/s, "
              f"stage-2 {roof and roofline_for(r)['achieved']} GB/s", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(dgc)
        print(f"# CPU baseline: {cpu['value'] / 1e6:.2f} M DOF-stage/s ({cpu['kind']})", file=sys.stderr)

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
                           autotune_stage2_us={k: round(v * 1e3, 2) for k, v in r["tune"]["ms"].items()})
                      for r in per],
        **extra,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
