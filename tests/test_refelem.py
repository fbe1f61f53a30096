"""Reference-triangle operators: product set-up vs the reference's own
matrices (golden vectors from the compiled reference, N = 1..6) and vs
identities the mathematics fixes (all N, including N = 7 which the
reference cannot build)."""
import numpy as np
import pytest

import dgc_amd

ORDERS = range(1, 8)


@pytest.mark.parametrize("N", range(1, 7))
def test_operators_match_reference(golden, N):
    g = golden("refelem.npz")
    w = dgc_amd.ref_operators(N, "weak")
    s = dgc_amd.ref_operators(N, "strong")
    np.testing.assert_allclose(w["r"], g[f"N{N}_r"], atol=1e-14, rtol=0)
    np.testing.assert_allclose(w["s"], g[f"N{N}_s"], atol=1e-14, rtol=0)
    for f in range(3):
        np.testing.assert_array_equal(w["fmask"][f], g[f"N{N}_f{f}"].astype(int))
    Np = w["Np"]
    scale = max(1.0, np.abs(g[f"N{N}_Dr_w"]).max())
    np.testing.assert_allclose(w["Dr"], g[f"N{N}_Dr_w"].reshape(Np, Np), atol=1e-12 * scale, rtol=0)
    np.testing.assert_allclose(w["Ds"], g[f"N{N}_Ds_w"].reshape(Np, Np), atol=1e-12 * scale, rtol=0)
    np.testing.assert_allclose(w["LIFT"], g[f"N{N}_L_w"].reshape(Np, -1), atol=1e-12 * scale, rtol=0)
    np.testing.assert_allclose(s["Dr"], g[f"N{N}_Dr_s"].reshape(Np, Np), atol=1e-12 * scale, rtol=0)
    np.testing.assert_allclose(s["Ds"], g[f"N{N}_Ds_s"].reshape(Np, Np), atol=1e-12 * scale, rtol=0)
    np.testing.assert_allclose(s["LIFT"], g[f"N{N}_L_s"].reshape(Np, -1), atol=1e-12 * scale, rtol=0)


# Closed-form Gauss-Lobatto points (the reference's src/quadrature_rules.h:80-145)
def _lgl_closed(npts):
    q = np.sqrt
    table = {
        2: [-1, 1],
        3: [-1, 0, 1],
        4: [-1, -q(1 / 5), q(1 / 5), 1],
        5: [-1, -q(3 / 7), 0, q(3 / 7), 1],
        6: [-1, -q(1 / 3 + 2 * q(7) / 21), -q(1 / 3 - 2 * q(7) / 21), q(1 / 3 - 2 * q(7) / 21),
            q(1 / 3 + 2 * q(7) / 21), 1],
        7: [-1, -q(5 / 11 + 2 * q(5 / 3) / 11), -q(5 / 11 - 2 * q(5 / 3) / 11), 0,
            q(5 / 11 - 2 * q(5 / 3) / 11), q(5 / 11 + 2 * q(5 / 3) / 11), 1],
    }
    return np.array(table[npts])


@pytest.mark.parametrize("N", range(1, 7))
def test_edge_nodes_are_lgl(N):
    w = dgc_amd.ref_operators(N, "weak")
    edge = w["r"][w["fmask"][0]]
    np.testing.assert_allclose(edge, _lgl_closed(N + 1), atol=2e-15)
    np.testing.assert_allclose(w["s"][w["fmask"][0]], -1.0, atol=2e-15)


@pytest.mark.parametrize("N", ORDERS)
def test_strong_derivative_polynomial_exactness(N):
    s = dgc_amd.ref_operators(N, "strong")
    r, ss = s["r"], s["s"]
    for a in range(N + 1):
        for b in range(N + 1 - a):
            f = r**a * ss**b
            dfr = a * r ** max(a - 1, 0) * ss**b if a else 0 * r
            dfs = b * r**a * ss ** max(b - 1, 0) if b else 0 * r
            np.testing.assert_allclose(s["Dr"] @ f, dfr, atol=1e-11 * N**2)
            np.testing.assert_allclose(s["Ds"] @ f, dfs, atol=1e-11 * N**2)


@pytest.mark.parametrize("N", ORDERS)
def test_weak_form_is_adjoint_of_strong(N):
    # Dr_w = V Vr^T M = M^{-1} Dr_s^T M
    w = dgc_amd.ref_operators(N, "weak")
    s = dgc_amd.ref_operators(N, "strong")
    M = w["M"]
    Minv = np.linalg.inv(M)
    scale = np.abs(w["Dr"]).max()
    np.testing.assert_allclose(w["Dr"], Minv @ s["Dr"].T @ M, atol=1e-11 * scale)
    np.testing.assert_allclose(w["Ds"], Minv @ s["Ds"].T @ M, atol=1e-11 * scale)


@pytest.mark.parametrize("N", ORDERS)
def test_mass_and_lift_identities(N):
    w = dgc_amd.ref_operators(N, "weak")
    M = w["M"]
    one = np.ones(w["Np"])
    assert abs(one @ M @ one - 2.0) < 1e-12  # area of the reference triangle
    # M LIFT restricted to face 0 nodes is the 1D LGL mass matrix
    E = M @ w["LIFT"]
    Nfp = w["Nfp"]
    f0 = w["fmask"][0]
    face_mass = E[np.ix_(f0, range(Nfp))]
    x = w["r"][f0]
    V1 = np.array([[_legendre_orthonormal(k, xi) for k in range(Nfp)] for xi in x])
    np.testing.assert_allclose(face_mass, np.linalg.inv(V1 @ V1.T), atol=1e-12)
    # SBP: S_r + S_r^T = E_r (boundary integral of phi_i phi_j n_r); check via 1^T
    s = dgc_amd.ref_operators(N, "strong")
    Sr = M @ s["Dr"]
    # integral of d/dr(phi_j) over the triangle = boundary integral of phi_j n_r
    np.testing.assert_allclose(one @ Sr @ one, 0.0, atol=1e-12)


def _legendre_orthonormal(k, x):
    from numpy.polynomial import legendre

    c = np.zeros(k + 1)
    c[k] = 1.0
    return legendre.legval(x, c) * np.sqrt((2 * k + 1) / 2.0)


def test_face_mass_order2_known_values():
    w = dgc_amd.ref_operators(2, "weak")
    E = w["M"] @ w["LIFT"]
    f0 = w["fmask"][0]
    np.testing.assert_allclose(E[np.ix_(f0, range(3))][0], [4 / 15, 2 / 15, -1 / 15], atol=1e-14)
