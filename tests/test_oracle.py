"""Pins the oracle (the CPU restatement of the reference's serial path) to
the golden vectors written by the compiled reference, and to properties the
mathematics fixes where the reference has no counterpart."""
import os
import numpy as np
import pytest

import dgc_amd
from oracle_lib import Problem


def _ref_ops(golden, N, form):
    g = golden("refelem.npz")
    Np = (N + 1) * (N + 2) // 2
    sfx = "w" if form == "weak" else "s"
    return dict(Dr=g[f"N{N}_Dr_{sfx}"].reshape(Np, Np), Ds=g[f"N{N}_Ds_{sfx}"].reshape(Np, Np),
                LIFT=g[f"N{N}_L_{sfx}"].reshape(Np, -1), r=g[f"N{N}_r"], s=g[f"N{N}_s"])


def _problem(golden, fix, form="weak", prefix=""):
    N = int(fix[prefix + "order"])
    m = dgc_amd.Mesh.from_arrays(fix[prefix + "vx"], fix[prefix + "vy"], fix[prefix + "etov"])
    return Problem.from_mesh(m, N, ops=_ref_ops(golden, N, form)), m


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_maxwell_rhs_matches_reference(golden):
    fix = golden("maxwell_c1.npz")
    P, _ = _problem(golden, fix)
    n = P.K * P.Np
    u = np.concatenate([np.zeros(2 * n), np.sin(np.pi * fix["x"]) * np.sin(np.pi * fix["y"])])
    assert rel(P.maxwell_rhs(u), fix["rhs0"]) < 1e-14


def test_maxwell_rk4_to_t1_matches_reference(golden):
    fix = golden("maxwell_c1.npz")
    P, _ = _problem(golden, fix)
    n = P.K * P.Np
    u0 = np.concatenate([np.zeros(2 * n), np.sin(np.pi * fix["x"]) * np.sin(np.pi * fix["y"])])
    u1, steps = P.rk4(0, u0, 0.0, float(fix["dt"]), nsteps=1)
    assert rel(u1, fix["step1"]) < 1e-14
    uT, steps = P.rk4(0, u0, 0.0, float(fix["dt"]), t_end=1.0)
    assert steps == int(fix["steps"])
    assert rel(uT, fix["final"]) < 1e-13
    # and the reference itself is close to the analytic cavity mode
    x, y, t = fix["x"], fix["y"], 1.0
    s2 = np.sqrt(2.0)
    exact = np.concatenate([-np.sin(np.pi * x) * np.cos(np.pi * y) * np.sin(s2 * np.pi * t) / s2,
                            np.cos(np.pi * x) * np.sin(np.pi * y) * np.sin(s2 * np.pi * t) / s2,
                            np.sin(np.pi * x) * np.sin(np.pi * y) * np.cos(s2 * np.pi * t)])
    assert np.mean((uT - exact) ** 2) < 1e-8


def test_advection_matches_reference(golden):
    fix = golden("advection_ref.npz")
    P, _ = _problem(golden, fix)
    assert rel(P.advection_rhs(fix["ic"], float(fix["t_rhs"]), 1.0, 1.0, 0), fix["rhs"]) < 1e-14
    u, _ = P.rk4(1, fix["ic"], 0.0, float(fix["dt"]), nsteps=int(fix["nsteps"]), ax=1.0, ay=1.0, bc=0)
    assert rel(u, fix["final"]) < 1e-14


def test_laplace_matches_reference(golden):
    fix = golden("laplace_ref.npz")
    P, _ = _problem(golden, fix, form="strong")
    for vin, vout in zip(fix["vin"], fix["vout"]):
        assert rel(P.laplace(vin, 0), vout) < 1e-14
    P2, _ = _problem(golden, fix, form="strong", prefix="coarse_")
    n = P2.K * P2.Np
    A = np.stack([P2.laplace(e, 0) for e in np.eye(n)], axis=1)
    np.testing.assert_allclose(A, fix["coarse_A"], atol=1e-12 * np.abs(A).max(), rtol=0)


def _mass_blocks(mesh, N):
    w = dgc_amd.ref_operators(N, "weak")
    J = mesh.geometry(N)["J"]
    return [Jc * w["M"] for Jc in J]


@pytest.mark.parametrize("tau_mode", [0, 1])
def test_laplace_minus_MA_symmetric_positive_definite(tau_mode):
    N = 2
    m = dgc_amd.Mesh.structured(4, 4, -1, 1, -1, 1, jitter=0.1, seed=9)
    P = Problem.from_mesh(m, N, form="strong")
    n = P.K * P.Np
    A = np.stack([P.laplace(e, tau_mode) for e in np.eye(n)], axis=1)
    from scipy.linalg import block_diag

    MA = -block_diag(*_mass_blocks(m, N)) @ A
    assert np.abs(MA - MA.T).max() < 1e-12 * np.abs(MA).max()
    ev = np.linalg.eigvalsh(0.5 * (MA + MA.T))
    assert ev.min() > 0


def test_laplace_planted_mistake_is_caught(golden):
    """A dropped penalty (tau = 0) must break the golden comparison; this
    guards the check itself against passing vacuously."""
    fix = golden("laplace_ref.npz")
    P, _ = _problem(golden, fix, form="strong")
    good = P.laplace(fix["vin"][0], 0)
    bad = P.laplace(fix["vin"][0], 1)  # different penalty: must not match
    assert rel(good, fix["vout"][0]) < 1e-14
    assert rel(bad, fix["vout"][0]) > 1e-3


@pytest.mark.parametrize("tau_mode", [0, 1])
def test_laplace_polynomial_exactness_in_interior(tau_mode):
    N = 3
    m = dgc_amd.Mesh.structured(8, 8, -1, 1, -1, 1, jitter=0.1, seed=4)
    P = Problem.from_mesh(m, N, form="strong")
    x, y = m.dof_coords(N, node_major=False)
    lap = P.laplace(x**2 + 3 * y**2 - x * y, tau_mode).reshape(P.K, P.Np)
    # cells two rings away from the boundary see no boundary penalty
    nb = P.nb_cell.reshape(-1, 3)
    bnd = {c for c in range(P.K) if (nb[c] == c).any()}
    ring = bnd | {int(o) for c in bnd for o in nb[c]}
    inner = [c for c in range(P.K) if c not in ring and not any(int(o) in ring for o in nb[c])]
    assert len(inner) > 10
    np.testing.assert_allclose(lap[inner], 8.0, atol=1e-9)


def test_maxwell_energy_nonincreasing_and_advection_conservation():
    N = 3
    w = dgc_amd.ref_operators(N, "weak")
    rng = np.random.default_rng(0)
    m = dgc_amd.Mesh.structured(6, 6, -1, 1, -1, 1, jitter=0.1, seed=2)
    P = Problem.from_mesh(m, N, ops=w)
    Mb = _mass_blocks(m, N)
    u = rng.standard_normal(3 * P.K * P.Np)
    du = P.maxwell_rhs(u)
    dE = sum(f.reshape(P.K, P.Np)[c] @ Mb[c] @ g.reshape(P.K, P.Np)[c]
             for f, g in zip(np.split(u, 3), np.split(du, 3)) for c in range(P.K))
    assert dE < 0  # upwind dissipation (and PEC is lossless)
    mp = dgc_amd.Mesh.structured(6, 6, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
    Pp = Problem.from_mesh(mp, N, ops=w)
    Mp = _mass_blocks(mp, N)
    v = rng.standard_normal(Pp.K * Pp.Np)
    dv = Pp.advection_rhs(v, 0.0, 1.0, 0.5, 1).reshape(Pp.K, Pp.Np)
    one = np.ones(Pp.Np)
    mass_rate = sum(one @ Mp[c] @ dv[c] for c in range(Pp.K))
    assert abs(mass_rate) < 1e-10 * np.abs(dv).max()
    vv = v.reshape(Pp.K, Pp.Np)
    assert sum(vv[c] @ Mp[c] @ dv[c] for c in range(Pp.K)) < 0
    const = Pp.advection_rhs(np.ones(Pp.K * Pp.Np), 0.0, 1.0, 0.5, 1)
    assert np.abs(const).max() < 1e-10


def test_threaded_oracle_gives_the_same_bits():
    """The all-core build of the restatement (OpenMP cell loops) changes no bit."""
    import dgc_amd
    from oracle_lib import Problem

    mesh = dgc_amd.Mesh.structured(12, 10, -1, 1, -1, 1, jitter=0.1, seed=4)
    for phys, N in ((0, 3), (1, 5)):
        P1 = Problem.from_mesh(mesh, N)
        P2 = Problem.from_mesh(mesh, N, threaded=True)
        F = 3 if phys == 0 else 1
        u0 = np.random.default_rng(N).standard_normal(F * P1.K * P1.Np)
        a, _ = P1.rk4(phys, u0, 0.0, 1e-3, nsteps=3, ax=1.0, ay=0.5, bc=1)
        b, _ = P2.rk4(phys, u0, 0.0, 1e-3, nsteps=3, ax=1.0, ay=0.5, bc=1)
        np.testing.assert_array_equal(a, b)


def _graddiv_case(golden, name):
    fix = golden("graddiv_ref.npz")
    P, m = _problem(golden, fix, prefix=name + "_")
    geo = m.geometry(P.N)
    bx, by = geo["bx"], geo["by"]
    # bc 1 of the fixture in affine form: exterior = g + beta * interior
    g = {0: (None, None, 1.0), 1: (2 * np.sin(bx) * np.cos(by), 2 * np.cos(bx) * np.sin(by), -1.0)}
    return fix, P, g


@pytest.mark.parametrize("name", ["coarse", "grid"])
def test_grad_and_div_match_reference(golden, name):
    fix, P, g = _graddiv_case(golden, name)
    for bc in (0, 1):
        g0, g1, beta = g[bc]
        for k in range(2):
            u0, u1 = fix[f"{name}_vin"][k]
            gx, gy = P.grad(u0, g0, beta)
            want = fix[f"{name}_grad_bc{bc}"][k]
            assert rel(gx, want[0]) < 1e-14 and rel(gy, want[1]) < 1e-14
            assert rel(P.div(u0, u1, g0, g1, beta), fix[f"{name}_div_bc{bc}"][k]) < 1e-14


def test_grad_planted_mistake_is_caught(golden):
    """Dropping the boundary data (g = 0 instead of 2 sin(x) cos(y)) must break
    the golden comparison."""
    fix, P, g = _graddiv_case(golden, "coarse")
    u0 = fix["coarse_vin"][0][0]
    gx, _ = P.grad(u0, None, -1.0)
    assert rel(gx, fix["coarse_grad_bc1"][0][0]) > 1e-3


@pytest.mark.parametrize("N", [1, 2, 3, 5])
def test_grad_and_div_exact_for_polynomials(N):
    """With exact boundary data the weak gradient of a polynomial of degree N
    is its exact gradient (the central flux reproduces the continuous trace),
    and likewise the divergence: a closed form the reference has no fixture for."""
    m = dgc_amd.Mesh.structured(6, 5, -1, 1, -1, 1, jitter=0.15, seed=9)
    P = Problem.from_mesh(m, N)
    x, y = m.dof_coords(N, node_major=False)
    geo = m.geometry(N)
    bx, by = geo["bx"], geo["by"]
    p = lambda x, y: x**N + 2 * x * y**(N - 1) - 3 * y**N + 0.5  # noqa: E731
    px = lambda x, y: N * x**(N - 1) + 2 * y**(N - 1)  # noqa: E731
    py = lambda x, y: 2 * (N - 1) * x * y**max(N - 2, 0) * (N > 1) - 3 * N * y**(N - 1)  # noqa: E731
    gx, gy = P.grad(p(x, y), 2 * p(bx, by), -1.0)
    np.testing.assert_allclose(gx, px(x, y), atol=1e-11)
    np.testing.assert_allclose(gy, py(x, y), atol=1e-11)
    q = lambda x, y: y**N - x * y**(N - 1)  # noqa: E731
    qy = N * y**(N - 1) - (N - 1) * x * y**max(N - 2, 0)
    d = P.div(p(x, y), q(x, y), 2 * p(bx, by), 2 * q(bx, by), -1.0)
    np.testing.assert_allclose(d, px(x, y) + qy, atol=1e-11)


def _laplace_bc_case(golden, name):
    fix = golden("laplace_bc_ref.npz")
    P, m = _problem(golden, fix, form="strong", prefix=name + "_")
    geo = m.geometry(P.N)
    bx, by = geo["bx"], geo["by"]
    gu = 2 * np.sin(bx) * np.cos(by)               # Dirichlet: exterior_val(x, y, 0)
    hq = 2 * (np.cos(bx) * np.sin(by) + bx * by)   # Neumann: exterior_grad_n(x, y, 0)
    return fix, P, gu, hq


@pytest.mark.parametrize("name", ["grid", "coarse"])
def test_laplace_boundary_kinds_match_reference(golden, name):
    fix, P, gu, hq = _laplace_bc_case(golden, name)
    zero = np.zeros(P.K * P.Np)
    for kind in (0, 1):
        inh = P.laplace_inhomogeneous(zero, kind, gu, hq)
        assert rel(inh, fix[f"{name}_inh_k{kind}"]) < 1e-14
        for k in range(2):
            hom = P.laplace_bc(fix[f"{name}_vin"][k], kind)
            assert rel(hom, fix[f"{name}_hom_k{kind}"][k]) < 1e-14
            # operator() = apply_homogeneous - inhomogeneous part (src/laplace_2d.h:121-128)
            assert rel(hom - inh, fix[f"{name}_full_k{kind}"][k]) < 1e-13


def test_laplace_boundary_kind_planted_mistake_is_caught(golden):
    fix, P, gu, hq = _laplace_bc_case(golden, "grid")
    wrong = P.laplace_bc(fix["grid_vin"][0], 0)  # Dirichlet where the fixture is Neumann
    assert rel(wrong, fix["grid_hom_k1"][0]) > 1e-3


def test_neumann_laplacian_annihilates_constants():
    """With Neumann faces everywhere the constants span the null space."""
    m = dgc_amd.Mesh.structured(7, 6, -1, 1, -1, 1, jitter=0.15, seed=2)
    for N in (1, 3, 5):
        P = Problem.from_mesh(m, N, form="strong")
        assert np.abs(P.laplace_bc(np.ones(P.K * P.Np), 1)).max() < 1e-10
        assert np.abs(P.laplace_bc(np.ones(P.K * P.Np), 0)).max() > 1.0


def test_laplace_data_without_data_is_the_homogeneous_operator():
    m = dgc_amd.Mesh.structured(5, 4, -1, 1, -1, 1, jitter=0.1, seed=1)
    P = Problem.from_mesh(m, 3, form="strong")
    u = np.random.default_rng(3).standard_normal(P.K * P.Np)
    for kind in (0, 1):
        assert np.array_equal(P.laplace_data(u, kind), P.laplace_bc(u, kind))


def _lap_errors(nxy, N=3):
    """Relative L2 errors of Lap_h u against the exact Laplacian of a smooth u
    with its Dirichlet data: the reference's operator()(u, bc) (homogeneous
    part minus apply_inhomogeneous_as_rhs) and the whole BR1 operator."""
    m = dgc_amd.Mesh.structured(nxy, nxy, -1, 1, -1, 1, jitter=0.1, seed=3)
    P = Problem.from_mesh(m, N, form="strong")
    x, y = m.dof_coords(N, node_major=False)
    geo = m.geometry(N)
    bx, by = geo["bx"], geo["by"]
    u = np.sin(x) * np.cosh(0.5 * y)
    lap = -0.75 * u
    g = 2 * np.sin(bx) * np.cosh(0.5 * by)
    ref = P.laplace_bc(u, 0) - P.laplace_inhomogeneous(np.zeros_like(u), 0, g, None)
    full = P.laplace_data(u, 0, g, None)
    return (np.linalg.norm(ref - lap) / np.linalg.norm(lap), np.linalg.norm(full - lap) / np.linalg.norm(lap))


def test_laplace_with_dirichlet_data_is_consistent_only_with_the_whole_affine_part():
    """The closed form pinning oracle_laplace_apply_data: the error of the whole
    BR1 operator with Dirichlet data falls as the mesh is refined, while the
    reference's operator()(u, bc), whose inhomogeneous part leaves out the
    interior faces of boundary cells, does not converge."""
    e4, e8, e16 = _lap_errors(4), _lap_errors(8), _lap_errors(16)
    assert e8[1] < e4[1] / 2.5 and e16[1] < e8[1] / 2.5 and e16[1] < 0.05
    assert e16[0] > e8[0] > e4[0] > 1.0


def test_stokes_oracle_converges_to_the_exact_solution():
    """The restated dual-splitting step (oracle/stokes_oracle.py) on the
    example's analytic problem: five steps of dt = 1e-3 from the exact data at
    t = 0 (the example's start: u^{n-1} = u^n), mean-square error as the
    example reports it, on 4 x 4 and 8 x 8 jittered meshes at N = 2."""
    import os
    import sys

    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    import stokes_oracle as so

    errs = []
    for nxy in (4, 8):
        m = dgc_amd.Mesh.structured(nxy, nxy, -1, 1, -1, 1, jitter=0.1, seed=3)
        N = 2
        ow, os_ = dgc_amd.ref_operators(N, "weak"), dgc_amd.ref_operators(N, "strong")
        Pw, Ps = Problem.from_mesh(m, N, ops=ow), Problem.from_mesh(m, N, ops=os_)
        c = m.cells()
        geo = so.Geometry(c["cx"], c["cy"], c["nb_cell"], os_["r"], os_["s"], os_["fmask"])
        x, y = m.dof_coords(N, node_major=False)
        up = u = so.exact(x, y, 0.0)
        tp = t = 0.0
        dt = 1e-3
        for _ in range(5):
            un = so.advance_timestep(Pw, Ps, geo, os_["M"], up, tp, u, t, dt)
            up, tp, u, t = u, t, un, t + dt
        ex = so.exact(x, y, t)
        errs.append((np.sum((u[0] - ex[0]) ** 2) + np.sum((u[1] - ex[1]) ** 2)) / len(x))
    assert errs[0] < 5e-3 and errs[1] < errs[0] / 10


# The reference's Poisson example (examples/poisson_2d/main.cpp:58-236: -2 pi^2
# sin(pi x) sin(pi y) as right-hand side of the collocated Laplacian with tau = 1,
# unit-vector assembly and a sparse direct solve; error norm = mean-square nodal
# error against sin(pi x) sin(pi y)) run on square_domain_fine.msh (K = 1064) by
# the compiled reference: error norm 3.3e-5 (N = 1) and 5.5e-8 (N = 2), max error
# 1.8e-2 and 1.4e-3, 23 and 40 nonzeros per row.
POISSON_FINE = {1: (3.3e-5, 1.8e-2, 23), 2: (5.5e-8, 1.4e-3, 40)}


@pytest.mark.parametrize("N", [1, 2])
def test_oracle_poisson_example_on_fine_mesh(N):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl

    mesh = dgc_amd.Mesh.read_gmsh(os.path.join(os.path.dirname(__file__), "golden", "square_domain_fine.msh"))
    P = Problem.from_mesh(mesh, N, form="strong")
    n = P.K * P.Np
    rows, cols, vals = [], [], []
    e = np.zeros(n)
    for i in range(n):
        e[i] = 1.0
        c = P.laplace(e, 0)
        e[i] = 0.0
        nz = np.nonzero(c)[0]
        rows.append(nz)
        cols.append(np.full(len(nz), i))
        vals.append(c[nz])
    A = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    x, y = mesh.dof_coords(N, node_major=False)
    exact = np.sin(np.pi * x) * np.sin(np.pi * y)
    u = spl.spsolve(A, -2 * np.pi**2 * exact)
    mse, mx, nnz = POISSON_FINE[N]
    assert abs(np.sum((u - exact) ** 2) / n - mse) <= 0.05 * mse
    assert abs(np.abs(u - exact).max() - mx) <= 0.05 * mx
    assert round(A.nnz / n) == nnz


# ---- the generalised advection operator (a != (1, 1), periodic) -------------
# The reference hard-codes a = (1, 1) (examples/advection_2d/advection_2d.h:216-218),
# so the oracle's general velocity is pinned by what the mathematics fixes:
# consistency with -a.grad u, the exact upwind energy identity, and the
# convergence of a periodic translation to the exact translated field.

def _periodic(n, N, jitter=0.05, ny=None):
    m = dgc_amd.Mesh.structured(n, ny or n, 0, 1, 0, 1, jitter=jitter, seed=1, periodic=True)
    return m, Problem.from_mesh(m, N)


def _sine_field(x, y):
    u = np.sin(2 * np.pi * x) * np.cos(4 * np.pi * y)
    ux = 2 * np.pi * np.cos(2 * np.pi * x) * np.cos(4 * np.pi * y)
    uy = -4 * np.pi * np.sin(2 * np.pi * x) * np.sin(4 * np.pi * y)
    return u, ux, uy


@pytest.mark.parametrize("a", [(1.0, 0.5), (-0.3, 0.8)])
def test_advection_rhs_is_minus_a_dot_grad(a):
    """On the periodic 24 x 24 mesh at N = 7 the RHS of a smooth periodic field
    equals -(a . grad u) to the discretization error (O(h^N): 5e-9 here, and
    1/2^N of the 12 x 12 error); a swapped (ax, ay) or a flipped ay is O(1) off."""
    ax, ay = a
    errs = []
    for n in (12, 24):
        m, P = _periodic(n, 7)
        x, y = m.dof_coords(7, node_major=False)
        u, ux, uy = _sine_field(x, y)
        exact = -(ax * ux + ay * uy)
        errs.append(rel(P.advection_rhs(u, 0.0, ax, ay, 1), exact))
        if n == 24:
            assert errs[-1] < 1e-7
            assert rel(P.advection_rhs(u, 0.0, ay, ax, 1), exact) > 0.5    # planted: swapped components
            assert rel(P.advection_rhs(u, 0.0, ax, -ay, 1), exact) > 0.5   # planted: wrong sign of ay
    assert errs[0] / errs[1] > 0.7 * 2**7


def _face_mass(xi):
    """1D mass matrix at the face nodes xi of [-1, 1]: (V V^T)^{-1} with the
    orthonormal Legendre Vandermonde."""
    from numpy.polynomial import legendre

    n = len(xi)
    V = np.stack([legendre.legval(xi, np.eye(n)[k]) * np.sqrt((2 * k + 1) / 2) for k in range(n)], axis=1)
    return np.linalg.inv(V @ V.T)


def _upwind_energy_rate(N, ax, ay, u, m, ops):
    """(d/dt of 1/2 |u|_M^2 under the oracle's RHS, the closed form
    -1/2 sum over faces of |a.n| int [u]^2 ds). For the weak form with upwind
    flux on a periodic mesh the two are equal identically in u."""
    P = Problem.from_mesh(m, N, ops=ops)
    K, Np = P.K, P.Np
    J = m.geometry(N)["J"]
    c = m.cells()
    cx, cy = c["cx"].reshape(K, 3), c["cy"].reshape(K, 3)
    nbc, nbf = c["nb_cell"].reshape(K, 3), c["nb_face"].reshape(K, 3)
    fm, r, s = ops["fmask"], ops["r"], ops["s"]
    M1 = [_face_mass(r[fm[0]]), _face_mass(s[fm[1]]), _face_mass(s[fm[2]])]
    uc = u.reshape(K, Np)
    duc = P.advection_rhs(u, 0.0, ax, ay, 1).reshape(K, Np)
    rate = sum(J[k] * uc[k] @ ops["M"] @ duc[k] for k in range(K))
    diss = 0.0
    for k in range(K):
        for e in range(3):
            dx, dy = cx[k, (e + 1) % 3] - cx[k, e], cy[k, (e + 1) % 3] - cy[k, e]
            ln = np.hypot(dx, dy)
            an = (ax * dy - ay * dx) / ln  # a . outward normal of a counter-clockwise cell
            jmp = uc[k, fm[e]] - uc[nbc[k, e], fm[nbf[k, e]][::-1]]
            diss += abs(an) * 0.5 * ln * (jmp @ M1[e] @ jmp)
    return rate, -0.25 * diss  # every face is visited from both sides


@pytest.mark.parametrize("N", [1, 3, 7])
@pytest.mark.parametrize("a", [(1.0, 0.5), (-0.3, 0.8)])
def test_advection_upwind_energy_identity(N, a):
    """The semi-discrete energy rate of a random field equals the upwind
    dissipation -1/2 sum |a.n| int [u]^2 to rounding; a central or downwind
    flux, a wrong flux scale or a wrong velocity breaks it."""
    m = dgc_amd.Mesh.structured(6, 5, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
    ops = dgc_amd.ref_operators(N, "weak")
    P = Problem.from_mesh(m, N, ops=ops)
    u = np.random.default_rng(N).standard_normal(P.K * P.Np)
    rate, closed = _upwind_energy_rate(N, a[0], a[1], u, m, ops)
    assert abs(rate - closed) < 1e-12 * abs(closed)
    # planted: the closed form of the swapped velocity is a different number
    _, swapped = _upwind_energy_rate(N, a[1], a[0], u, m, ops)
    assert abs(rate - swapped) > 1e-3 * abs(closed)


def _translation_error(n, N, ax, ay, T=0.25, ex_a=None):
    m, P = _periodic(n, N)
    x, y = m.dof_coords(N, node_major=False)
    f = lambda x, y: np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)  # noqa: E731
    dt0 = 0.25 * m.min_edge / (np.hypot(ax, ay) * (N + 1) ** 2)
    ns = int(np.ceil(T / dt0))
    u, _ = P.rk4(1, f(x, y), 0.0, T / ns, nsteps=ns, ax=ax, ay=ay, bc=1)
    bx, by = ex_a if ex_a is not None else (ax, ay)
    return np.sqrt(np.mean((u - f(x - bx * T, y - by * T)) ** 2))


@pytest.mark.parametrize("a", [(1.0, 0.5), (-0.3, 0.8)])
def test_periodic_translation_converges_at_order_N_plus_1(a):
    """RK4 of the oracle on periodic meshes translates a periodic sine by a T:
    the RMS error against the exact translate falls by about 2^(N+1) per
    halving of h (N = 2: 8x), while the translate by the swapped velocity
    stays O(1) away."""
    e = [_translation_error(n, 2, *a) for n in (8, 16, 32)]
    assert e[0] / e[1] > 6.0 and e[1] / e[2] > 6.0 and e[2] < 3e-4
    assert _translation_error(16, 2, *a, ex_a=(a[1], a[0])) > 0.05
