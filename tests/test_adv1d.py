"""1D advection (examples/advection_1d): the oracle against the compiled
reference's golden vectors and the exact solution, the library's operators
against the reference's, and the examples' plot files (no GPU needed)."""
import numpy as np
import pytest

import dgc_amd
from oracle_lib import adv1d_rhs, adv1d_rk4

CASES = [(16, 3), (64, 6)]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("K,N", CASES)
def test_oracle_matches_reference(golden, K, N):
    z = golden("adv1d_ref.npz")
    p = f"K{K}_N{N}_"
    D, L = z[p + "D"], z[p + "L"]
    assert rel(adv1d_rhs(K, N, D, L, z[p + "ic"], 0.3), z[p + "rhs03"]) < 1e-14
    h = 2 * np.pi / K
    dt = 0.25 / N / N * h / (2 * np.pi)
    assert rel(adv1d_rk4(K, N, D, L, z[p + "ic"], 0.0, dt, int(z[p + "steps"])), z[p + "final"]) < 1e-13
    # a planted mistake (half the step) must fail
    assert rel(adv1d_rk4(K, N, D, L, z[p + "ic"], 0.0, dt / 2, int(z[p + "steps"])), z[p + "final"]) > 1e-6


@pytest.mark.parametrize("K,N", CASES)
def test_library_operators_match_reference(golden, K, N):
    z = golden("adv1d_ref.npz")
    p = f"K{K}_N{N}_"
    D, L, x = dgc_amd.Advection1D.operators(K, N)
    assert rel(D, z[p + "D"]) < 1e-13 and rel(L, z[p + "L"]) < 1e-13
    # node-major here, cell-major in the reference's CPU layout
    np.testing.assert_allclose(dgc_amd.node_to_cell_major(x, K, N + 1), z[p + "x"], rtol=0, atol=1e-14)


def test_oracle_converges_to_the_exact_solution():
    """Transport of sin(x): after 200 steps the error falls with the order."""
    errs = []
    for N in (2, 4):
        K = 32
        D, L, x = dgc_amd.Advection1D.operators(K, N)
        xc = dgc_amd.node_to_cell_major(x, K, N + 1)
        h = 2 * np.pi / K
        dt = 0.25 / N / N * h / (2 * np.pi)
        u = adv1d_rk4(K, N, D, L, np.sin(xc), 0.0, dt, 200)
        errs.append(np.mean((u - np.sin(xc - 2 * np.pi * 200 * dt)) ** 2))
    assert errs[1] < errs[0] / 100 and errs[1] < 1e-12


def test_plot_files_have_the_examples_format(tmp_path):
    x = np.array([0.0, 0.5, 1.0 / 3.0])
    y = np.array([1.0, -2.5e-7, 3.0])
    u = np.array([np.pi, 1e20, -0.0])
    ref = np.array([2.0, np.e, 1.0])
    p2 = tmp_path / "Advection2DDataFile.txt"
    dgc_amd.write_plot_file(p2, x, u, ref, y=y)
    f = lambda v: f"{v:.15g}"  # noqa: E731  (ostream, precision digits10 = 15)
    want2 = (["#         x         y         u"] + [f"{f(a)}  {f(b)}  {f(c)}" for a, b, c in zip(x, y, u)] + [""]
             + ["#         x         y         reference solution"]
             + [f"{f(a)}  {f(b)}  {f(c)}" for a, b, c in zip(x, y, ref)])
    assert p2.read_text().splitlines() == want2
    p1 = tmp_path / "Advection1DDataFile.txt"
    dgc_amd.write_plot_file(p1, x, u, ref)
    want1 = (["#         x         y"] + [f"{f(a)}  {f(c)}" for a, c in zip(x, u)] + [""]
             + ["#         x         reference solution"] + [f"{f(a)} {f(c)}" for a, c in zip(x, ref)])
    assert p1.read_text().splitlines() == want1
