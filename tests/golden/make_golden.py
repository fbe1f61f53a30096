"""Write the golden vectors under tests/golden/ from the compiled reference.

Runs in the build container only (it needs oracle/_ref/ref_harness, built by
oracle/build_ref.sh from the reference sources). The meshes are produced by
this project's generator and stored with the vectors, so the tests can feed
the very same mesh to the HIP path and to the oracle. Each file holds inputs
and the reference's outputs only (no reference source).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "dg-on-cuda_amd"))
import dgc_amd as dgc  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def write_mesh(path, vx, vy, etov):
    with open(path, "wb") as f:
        np.array([len(vx), len(etov)], dtype=np.int32).tofile(f)
        np.asarray(vx, np.float64).tofile(f)
        np.asarray(vy, np.float64).tofile(f)
        np.asarray(etov, np.int32).reshape(-1).tofile(f)


def run(*args):
    subprocess.run([HARNESS, *map(str, args)], check=True)


def mesh_arrays(mesh):
    vx, vy, etov = mesh.vertices()
    return vx, vy, etov


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build oracle/_ref first: oracle/build_ref.sh")
    tmp = tempfile.mkdtemp(prefix="dgc_golden_")
    out = {}

    # 1. reference-element operators N = 1..6 (the reference throws for N = 7)
    coarse = dgc.Mesh.read_gmsh(os.path.join(HERE, "square_domain_coarse.msh"))
    vx, vy, etov = mesh_arrays(coarse)
    mpath = os.path.join(tmp, "coarse.bin")
    write_mesh(mpath, vx, vy, etov)
    ref = {}
    for N in range(1, 7):
        Np, Nfp = (N + 1) * (N + 2) // 2, N + 1
        o = os.path.join(tmp, f"ref{N}.bin")
        run("refelem", mpath, N, o)
        a = np.fromfile(o)
        sizes = [Np, Np, Nfp, Nfp, Nfp, Np * Np, Np * Np, Np * 3 * Nfp, Np * Np, Np * Np, Np * 3 * Nfp]
        parts = np.split(a, np.cumsum(sizes)[:-1])
        names = ["r", "s", "f0", "f1", "f2", "Dr_w", "Ds_w", "L_w", "Dr_s", "Ds_s", "L_s"]
        for nm, p in zip(names, parts):
            ref[f"N{N}_{nm}"] = p
    np.savez_compressed(os.path.join(HERE, "refelem.npz"), **ref)

    # 2. Maxwell TM cavity: 16x16 squares on [-1,1]^2, jitter 0.1h, seed 0, N = 4,
    #    dt = 0.5 h_min / (N+1)^2 to t = 1 (truncated last step)
    N = 4
    m = dgc.Mesh.structured(16, 16, -1, 1, -1, 1, jitter=0.1, seed=0)
    vx, vy, etov = mesh_arrays(m)
    mpath = os.path.join(tmp, "c1.bin")
    write_mesh(mpath, vx, vy, etov)
    dt = 0.5 * m.min_edge / (N + 1) ** 2
    o = os.path.join(tmp, "geo.bin")
    run("geometry", mpath, N, o)
    K = m.K
    a = np.fromfile(o)
    sizes = [4 * K, K, 3 * K, 3 * K, 3 * K, 3 * K, 3 * K, 1]
    parts = np.split(a[: sum(sizes)], np.cumsum(sizes)[:-1])
    nb = int(parts[-1][0])
    rest = a[sum(sizes):]
    geo = dict(inv_jac=parts[0], J=parts[1], face_J=parts[2], iface_cell=parts[3].astype(np.int32),
               iface_face=parts[4].astype(np.int32), nx=parts[5], ny=parts[6],
               bx=rest[:nb], by=rest[nb:2 * nb])
    o = os.path.join(tmp, "mx.bin")
    run("maxwell", mpath, N, 1.0, repr(dt), o)
    Np = (N + 1) * (N + 2) // 2
    n = K * Np
    a = np.fromfile(o)
    xs, ys, ic, rhs, step1, final = np.split(a[: 14 * n], np.cumsum([n, n, 3 * n, 3 * n, 3 * n])[:]) [:6]
    steps, tfin, secs = a[14 * n:]
    np.savez_compressed(os.path.join(HERE, "maxwell_c1.npz"), vx=vx, vy=vy, etov=etov, order=N, dt=dt,
                        T=1.0, steps=steps, t_final=tfin, x=xs, y=ys, rhs0=rhs, step1=step1,
                        final=final, ref_seconds=secs, **{f"geo_{k}": v for k, v in geo.items()})
    print(f"maxwell C1: {int(steps)} steps, t = {tfin}, reference CPU {secs:.2f} s")

    # 3. reference advection problem (a = (1,1), analytic inflow): 16x16, jitter 0.1h, seed 5, N = 3
    N = 3
    m = dgc.Mesh.structured(16, 16, -1, 1, -1, 1, jitter=0.1, seed=5)
    vx, vy, etov = mesh_arrays(m)
    mpath = os.path.join(tmp, "adv.bin")
    write_mesh(mpath, vx, vy, etov)
    dt = 0.5 * m.min_edge / (N + 1) ** 2
    o = os.path.join(tmp, "adv_out.bin")
    run("advection", mpath, N, 0.3, 10, repr(dt), o)
    n = m.K * (N + 1) * (N + 2) // 2
    ic, rhs, fin = np.split(np.fromfile(o), 3)
    np.savez_compressed(os.path.join(HERE, "advection_ref.npz"), vx=vx, vy=vy, etov=etov, order=N,
                        dt=dt, t_rhs=0.3, nsteps=10, ic=ic, rhs=rhs, final=fin)

    # 4. laplace_2d (tau = 1, homogeneous Dirichlet): 16x16, jitter 0.1h, seed 3, N = 3,
    #    two seeded vectors; plus the dense operator on the K = 16 gmsh mesh at N = 2
    N = 3
    m = dgc.Mesh.structured(16, 16, -1, 1, -1, 1, jitter=0.1, seed=3)
    vx, vy, etov = mesh_arrays(m)
    mpath = os.path.join(tmp, "lap.bin")
    write_mesh(mpath, vx, vy, etov)
    n = m.K * (N + 1) * (N + 2) // 2
    rng = np.random.default_rng(7)
    vin = rng.standard_normal((2, n))
    ipath, opath = os.path.join(tmp, "lin.bin"), os.path.join(tmp, "lout.bin")
    vin.tofile(ipath)
    run("laplace", mpath, N, ipath, opath)
    vout = np.fromfile(opath).reshape(2, n)
    cvx, cvy, cetov = mesh_arrays(coarse)
    cpath = os.path.join(tmp, "coarse2.bin")
    write_mesh(cpath, cvx, cvy, cetov)
    n2 = coarse.K * 6
    np.eye(n2).tofile(ipath)
    run("laplace", cpath, 2, ipath, opath)
    A = np.fromfile(opath).reshape(n2, n2).T  # column i = A e_i
    np.savez_compressed(os.path.join(HERE, "laplace_ref.npz"), vx=vx, vy=vy, etov=etov, order=N,
                        vin=vin, vout=vout, coarse_vx=cvx, coarse_vy=cvy, coarse_etov=cetov,
                        coarse_order=2, coarse_A=A)
    graddiv(tmp)
    laplace_bc(tmp)
    adv1d(tmp)
    print("golden vectors written to", HERE)


def graddiv(tmp):
    """5. grad_2d and div_2d (central fluxes, weak form) on the K = 16 gmsh mesh at
    N = 3 and a jittered 8 x 6 mesh of [-1, 1]^2 at N = 2, two seeded vectors
    (pairs for div), BC 0 (exterior = interior) and BC 1 (exterior =
    2 sin(x) cos(y) - interior, second component 2 cos(x) sin(y) - interior)."""
    out = {}
    meshes = {"coarse": (dgc.Mesh.read_gmsh(os.path.join(HERE, "square_domain_coarse.msh")), 3),
              "grid": (dgc.Mesh.structured(8, 6, -1, 1, -1, 1, jitter=0.15, seed=11), 2)}
    for name, (m, N) in meshes.items():
        vx, vy, etov = mesh_arrays(m)
        mpath = os.path.join(tmp, f"gd_{name}.bin")
        write_mesh(mpath, vx, vy, etov)
        n = m.K * (N + 1) * (N + 2) // 2
        rng = np.random.default_rng(13)
        vin = rng.standard_normal((2, 2, n))  # two pairs
        ipath, opath = os.path.join(tmp, "gin.bin"), os.path.join(tmp, "gout.bin")
        out.update({f"{name}_vx": vx, f"{name}_vy": vy, f"{name}_etov": etov, f"{name}_order": N,
                    f"{name}_vin": vin})
        for bc in (0, 1):
            vin[:, 0].tofile(ipath)
            run("grad", mpath, N, bc, ipath, opath)
            out[f"{name}_grad_bc{bc}"] = np.fromfile(opath).reshape(2, 2, n)
            vin.tofile(ipath)
            run("div", mpath, N, bc, ipath, opath)
            out[f"{name}_div_bc{bc}"] = np.fromfile(opath).reshape(2, n)
    np.savez_compressed(os.path.join(HERE, "graddiv_ref.npz"), **out)


def laplace_bc(tmp):
    """6. laplace_2d with a BC object of one kind on every boundary node (kind 0
    Dirichlet, exterior_val = 2 sin(x) cos(y) - a; kind 1 Neumann,
    exterior_grad_n = 2 (cos(x) sin(y) + x y) - g): apply_homogeneous of two seeded
    vectors, apply_inhomogeneous_as_rhs of zero, operator() of the two vectors; on
    a jittered 8 x 6 mesh of [-1, 1]^2 at N = 3 and the K = 16 gmsh mesh at N = 2."""
    out = {}
    meshes = {"grid": (dgc.Mesh.structured(8, 6, -1, 1, -1, 1, jitter=0.15, seed=12), 3),
              "coarse": (dgc.Mesh.read_gmsh(os.path.join(HERE, "square_domain_coarse.msh")), 2)}
    for name, (m, N) in meshes.items():
        vx, vy, etov = mesh_arrays(m)
        mpath = os.path.join(tmp, f"lb_{name}.bin")
        write_mesh(mpath, vx, vy, etov)
        n = m.K * (N + 1) * (N + 2) // 2
        vin = np.random.default_rng(17).standard_normal((2, n))
        ipath, opath = os.path.join(tmp, "lbin.bin"), os.path.join(tmp, "lbout.bin")
        vin.tofile(ipath)
        out.update({f"{name}_vx": vx, f"{name}_vy": vy, f"{name}_etov": etov, f"{name}_order": N,
                    f"{name}_vin": vin})
        for kind in (0, 1):
            run("laplace_bc", mpath, N, kind, ipath, opath)
            r = np.fromfile(opath).reshape(5, n)
            out[f"{name}_hom_k{kind}"] = r[:2]
            out[f"{name}_inh_k{kind}"] = r[2]
            out[f"{name}_full_k{kind}"] = r[3:]
    np.savez_compressed(os.path.join(HERE, "laplace_bc_ref.npz"), **out)


def adv1d(tmp):
    """7. examples/advection_1d: (K, N, steps) = (16, 3, 50) and (64, 6, 20):
    the operators, node positions, initial state, its right-hand side at
    t = 0.3 and the state after the steps (dt = 0.25 / N^2 h / a)."""
    out = {}
    for K, N, steps in ((16, 3, 50), (64, 6, 20)):
        opath = os.path.join(tmp, "a1.bin")
        run("adv1d", K, N, steps, opath)
        Np, n = N + 1, K * (N + 1)
        r = np.fromfile(opath)
        D, L, rest = r[:Np * Np], r[Np * Np:2 * Np * Np], r[2 * Np * Np:].reshape(4, n)
        pre = f"K{K}_N{N}_"
        out.update({pre + "D": D.reshape(Np, Np), pre + "L": L.reshape(Np, Np), pre + "x": rest[0],
                    pre + "ic": rest[1], pre + "rhs03": rest[2], pre + "final": rest[3], pre + "steps": steps})
    np.savez_compressed(os.path.join(HERE, "adv1d_ref.npz"), **out)


if __name__ == "__main__":
    if sys.argv[1:] == ["graddiv"]:
        graddiv(tempfile.mkdtemp(prefix="dgc_golden_"))
    elif sys.argv[1:] == ["laplace_bc"]:
        laplace_bc(tempfile.mkdtemp(prefix="dgc_golden_"))
    elif sys.argv[1:] == ["adv1d"]:
        adv1d(tempfile.mkdtemp(prefix="dgc_golden_"))
    else:
        main()
