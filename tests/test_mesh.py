"""Mesh generation, topology and geometry (host set-up) vs the reference's
fill_* outputs on the same mesh, plus structural invariants."""
import os

import numpy as np
import pytest

import dgc_amd
from conftest import GOLDEN


def test_generator_is_deterministic_and_matches_fixture(golden):
    g = golden("maxwell_c1.npz")
    m = dgc_amd.Mesh.structured(16, 16, -1, 1, -1, 1, jitter=0.1, seed=0)
    vx, vy, etov = m.vertices()
    np.testing.assert_array_equal(vx, g["vx"])
    np.testing.assert_array_equal(vy, g["vy"])
    np.testing.assert_array_equal(etov, g["etov"])
    # jitter stays within +-0.1 h of the lattice and boundary vertices stay on the boundary
    lat = np.linspace(-1, 1, 17)
    X, Y = np.meshgrid(lat, lat)
    assert np.abs(vx - X.ravel()).max() <= 0.1 * 0.125 + 1e-15
    assert np.abs(vy - Y.ravel()).max() <= 0.1 * 0.125 + 1e-15
    onb = (np.abs(np.abs(X.ravel()) - 1) < 1e-14) | (np.abs(np.abs(Y.ravel()) - 1) < 1e-14)
    np.testing.assert_array_equal(vx[onb], X.ravel()[onb])


def test_geometry_matches_reference(golden):
    g = golden("maxwell_c1.npz")
    m = dgc_amd.Mesh.from_arrays(g["vx"], g["vy"], g["etov"])
    geo = m.geometry(int(g["order"]))
    for k in ("inv_jac", "J", "face_J", "nx", "ny", "bx", "by"):
        # closed-form 2x2 inverse vs Gauss-Jordan: rounding-level differences only
        scale = np.abs(g[f"geo_{k}"]).max()
        np.testing.assert_allclose(geo[k], g[f"geo_{k}"], rtol=0, atol=4e-16 * 16 * scale, err_msg=k)
    np.testing.assert_array_equal(geo["iface_cell"], g["geo_iface_cell"])
    np.testing.assert_array_equal(geo["iface_face"], g["geo_iface_face"])


def test_topology_is_symmetric():
    m = dgc_amd.Mesh.structured(9, 7, 0, 1, 0, 1, jitter=0.15, seed=11)
    c = m.cells()
    nc, nf = c["nb_cell"].reshape(-1, 3), c["nb_face"].reshape(-1, 3)
    nb = 0
    for cell in range(m.K):
        for e in range(3):
            o, f = nc[cell, e], nf[cell, e]
            if o == cell:
                nb += 1
                continue
            assert nc[o, f] == cell and nf[o, f] == e
    assert nb == 2 * (9 + 7)


def test_periodic_mesh_has_no_boundary():
    m = dgc_amd.Mesh.structured(8, 8, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True)
    c = m.cells()
    cells = np.repeat(np.arange(m.K), 3)
    assert np.all(c["nb_cell"] != cells)
    g = m.geometry(2)
    assert g["num_boundary_nodes"] == 0
    # twin faces: equal length, opposite normals
    fJ = g["face_J"].reshape(-1, 3)
    nx, ny = g["nx"].reshape(-1, 3), g["ny"].reshape(-1, 3)
    nc, nf = c["nb_cell"].reshape(-1, 3), c["nb_face"].reshape(-1, 3)
    for cell in range(m.K):
        for e in range(3):
            o, f = nc[cell, e], nf[cell, e]
            assert abs(fJ[cell, e] - fJ[o, f]) < 1e-14
            assert abs(nx[cell, e] + nx[o, f]) < 1e-13 and abs(ny[cell, e] + ny[o, f]) < 1e-13


def test_cells_are_counter_clockwise_and_tile_domain():
    m = dgc_amd.Mesh.structured(16, 16, -1, 1, -1, 1, jitter=0.1, seed=3)
    g = m.geometry(1)
    assert np.all(g["J"] > 0)
    assert abs(g["J"].sum() * 2 - 4.0) < 1e-12  # sum of areas = |[-1,1]^2|


def test_gmsh_reader_on_reference_meshes():
    m = dgc_amd.Mesh.read_gmsh(os.path.join(GOLDEN, "square_domain_coarse.msh"))
    assert m.K == 16
    g = m.geometry(2)
    assert np.all(g["J"] > 0)
    assert abs(g["J"].sum() * 2 - 4.0) < 1e-10
    m4 = dgc_amd.Mesh.read_gmsh(os.path.join(GOLDEN, "square_domain_4cells.msh"))
    assert m4.K == 4


def test_bad_arguments_raise():
    with pytest.raises(dgc_amd.DGCError):
        dgc_amd.Mesh.structured(0, 4)
    with pytest.raises(dgc_amd.DGCError):
        dgc_amd.Mesh.structured(4, 4, jitter=0.5)
    with pytest.raises(dgc_amd.DGCError):
        dgc_amd.Mesh.read_gmsh("/nonexistent.msh")
    with pytest.raises(dgc_amd.DGCError):
        dgc_amd.ref_operators(0)
