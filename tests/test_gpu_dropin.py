"""The C++ drop-in layer (dg-on-cuda_amd/host/dgc/*.h) end to end: the example
drivers written against it run on the GPU and reproduce the analytic
solutions, as the reference's example programs do."""
import os
import re
import subprocess

import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "dg-on-cuda_amd", "bin")


def run(exe, *args):
    out = subprocess.run([os.path.join(BIN, exe), *map(str, args)], capture_output=True, text=True, timeout=300)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stderr
    return out.stdout


def test_maxwell_driver_cavity_mode():
    out = run("maxwell_2d", 4, "structured")
    m = re.search(r"steps = (\d+).*error norm = ([0-9.e+-]+)", out)
    assert m and int(m.group(1)) == 491
    assert float(m.group(2)) < 1e-8


def test_maxwell_driver_with_device_setup():
    # geometry and connectivity built on the GPU: the same run to the last digit
    host = run("maxwell_2d", 4, "structured", "host")
    dev = run("maxwell_2d", 4, "structured", "device")
    pick = lambda out: re.search(r"steps = \d+.*error norm = [0-9.e+-]+", out).group(0)
    assert pick(dev) == pick(host)


def test_maxwell_driver_on_reference_gmsh_mesh():
    # K = 16 gmsh mesh of the reference examples; the reference CPU path gives
    # a mean-square error of about 1.7e-3 at N = 3, T = 1 on it
    out = run("maxwell_2d", 3, os.path.join(GOLDEN, "square_domain_coarse.msh"))
    m = re.search(r"error norm = ([0-9.e+-]+)", out)
    assert m and float(m.group(1)) < 5e-3


@pytest.mark.parametrize("tau_mode", [0, 1])
def test_poisson_driver(tau_mode):
    out = run("poisson_2d", 3, "structured", tau_mode)
    m = re.search(r"max error = ([0-9.e+-]+)", out)
    assert m and float(m.group(1)) < 1e-3


@pytest.mark.parametrize("order", [2, 3, 6])
def test_graddiv_driver_exact_for_polynomials(order):
    # dgc::grad_2d / dgc::div_2d with reference-style BC objects (Dirichlet data)
    out = run("graddiv_2d", order)
    m = re.search(r"grad max error = ([0-9.e+-]+), div max error = ([0-9.e+-]+)", out)
    assert m and float(m.group(1)) < 1e-9 and float(m.group(2)) < 1e-9


@pytest.mark.parametrize("order", [2, 4])
def test_laplace_driver_with_bc_objects(order):
    # dgc::laplace_2d::operator()(u, bc) with Dirichlet and Neumann BC objects
    # against the oracle (apply_homogeneous minus the inhomogeneous part)
    import numpy as np

    import dgc_amd
    from oracle_lib import Problem

    out = run("laplace_bc_2d", order)
    m = re.search(r"dirichlet norm = ([0-9.e+-]+), neumann norm = ([0-9.e+-]+)", out)
    mesh = dgc_amd.Mesh.structured(10, 8, -1, 1, -1, 1, jitter=0.15, seed=7)
    P = Problem.from_mesh(mesh, order, form="strong")
    x, y = mesh.dof_coords(order, node_major=False)
    u = np.sin(x) * np.cos(y)
    geo = mesh.geometry(order)
    bx, by = geo["bx"], geo["by"]
    gu = 2 * np.sin(bx) * np.cos(by)
    hq = 2 * (np.cos(bx) * np.sin(by) + bx * by)
    zero = np.zeros_like(u)
    for kind, got in ((0, float(m.group(1))), (1, float(m.group(2)))):
        want = np.linalg.norm(P.laplace_bc(u, kind) - P.laplace_inhomogeneous(zero, kind, gu, hq))
        assert abs(got - want) <= 1e-12 * want
    m = re.search(r"assembled: nnz = (\d+), max \|A u - Lap u\| / max \|Lap u\| = ([0-9.e+-]+)", out)
    assert m and int(m.group(1)) > 0 and float(m.group(2)) < 1e-13


def test_advection_1d_driver(tmp_path):
    # the 1D example's flow (examples/advection_1d/main.cpp) with its plot file
    plot = tmp_path / "Advection1DDataFile.txt"
    out = run("advection_1d", 512, 4, 2000, plot)
    m = re.search(r"error norm = ([0-9.e+-]+)", out)
    assert m and float(m.group(1)) < 1e-12
    lines = plot.read_text().splitlines()
    assert lines[0] == "#         x         y" and lines[512 * 5 + 2] == "#         x         reference solution"


@pytest.mark.parametrize("order,want", [(1, 3.3403e-05), (2, 5.4973e-08), (3, 3.6416e-11)])
def test_poisson_driver_reproduces_reference_error_on_fine_mesh(order, want):
    # The reference's Poisson example on square_domain_fine.msh (tau = 1): mean-square
    # error 3.3e-5, 5.5e-8 and 3.6e-11 at N = 1, 2, 3 from its direct solve (the
    # oracle's direct solve gives the 5 digits used here, tests/test_oracle.py);
    # the driver's PCG solve to 1e-10 reaches the same discrete solution.
    out = run("poisson_2d", order, os.path.join(GOLDEN, "square_domain_fine.msh"), 0)
    m = re.search(r"error norm = ([0-9.e+-]+)", out)
    assert m and abs(float(m.group(1)) - want) <= 1e-3 * want


@pytest.mark.parametrize("order,boxes", [(1, 0), (2, 8)])
def test_stokes_driver_on_fine_mesh(order, boxes):
    # the Stokes example's run (examples/stokes_2d/main.cpp:58-196) through
    # dgc_stokes_advance: 100 steps to T = 0.1 on the reference's fine mesh; the
    # error norms of the Python path (tests/test_gpu_stokes.py): 2.535e-5 at
    # N = 1, 4.48e-8 at N = 2 (the latter with two-level pressure solves)
    out = run("stokes_2d", order, os.path.join(GOLDEN, "square_domain_fine.msh"), boxes)
    m = re.search(r"T = ([0-9.e+-]+), error norm = ([0-9.e+-]+)", out)
    assert m and abs(float(m.group(1)) - 0.1) < 1e-12
    want = {1: 2.535e-5, 2: 4.477e-8}[order]
    assert abs(float(m.group(2)) - want) <= 0.01 * want
    assert "100 steps" in out


@pytest.mark.parametrize("order", [1, 2, 4])
def test_advection_2d_driver_matches_oracle_run(order, tmp_path):
    # the 2D advection example's run (examples/advection_2d/main.cpp:58-190): the
    # reference's coarse mesh (K = 16), 10000 RK4 steps of its dt; the error norm
    # equals the oracle's for the same run (the reference's CPU algorithm)
    import numpy as np

    import dgc_amd
    from oracle_lib import Problem

    path = os.path.join(GOLDEN, "square_domain_coarse.msh")
    plot = tmp_path / "Advection2DDataFile.txt"
    out = run("advection_2d", order, path, 10000, plot)
    m = re.search(r"T = ([0-9.e+-]+), error norm = ([0-9.e+-]+)", out)
    mesh = dgc_amd.Mesh.read_gmsh(path)
    P = Problem.from_mesh(mesh, order)
    x, y = mesh.dof_coords(order, node_major=False)
    l = 2 / np.sqrt(mesh.K)
    dt = (2 / 3) * (l / order) * (l / (2 + np.sqrt(2)))
    u, _ = P.rk4(1, np.sin((x + 1) * np.pi) * np.sin((y + 1) * np.pi), 0.0, dt, nsteps=10000)
    t = 0.0
    for _ in range(10000):
        t += dt
    ref = np.sin((x + 1 - t) * np.pi) * np.sin((y + 1 - t) * np.pi)
    want = np.sum((u - ref) ** 2) / len(x)
    assert m and abs(float(m.group(2)) - want) <= 1e-6 * want
    lines = plot.read_text().splitlines()
    n = mesh.K * (order + 1) * (order + 2) // 2
    assert lines[0] == "#         x         y         u" and len(lines) == 2 * n + 3
