"""The IPDG Laplacian assembled as a sparse matrix on the GPU (dgc_ipdg_assemble)
against the compiled reference's unit-vector assembly (tests/golden/laplace_ref.npz,
coarse_A: the dense operator of the K = 16 gmsh mesh at N = 2) and against the
operator's own application."""
import numpy as np
import pytest
import scipy.sparse as sp

import dgc_amd
from dgc_amd import cell_to_node_major as c2n

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_assembly_matches_reference_unit_vector_assembly(ctx, golden):
    fix = golden("laplace_ref.npz")
    mesh = dgc_amd.Mesh.from_arrays(fix["coarse_vx"], fix["coarse_vy"], fix["coarse_etov"])
    N = int(fix["coarse_order"])
    disc = dgc_amd.Discretization(ctx, mesh, N)
    ip = dgc_amd.IPDG(disc, tau_mode=0)
    rp, ci, va = ip.assemble()
    n, K, Np = disc.size, disc.K, disc.Np
    A = sp.csr_matrix((va, ci, rp), shape=(n, n)).toarray()
    # the fixture is in the reference CPU layout (cell-major c Np + j)
    perm = c2n(np.arange(n), K, Np)  # node-major position -> cell-major index
    want = fix["coarse_A"][np.ix_(perm, perm)]
    assert np.abs(A - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("N", [1, 3, 5])
@pytest.mark.parametrize("kind", [0, 1])
def test_assembly_reproduces_the_apply(ctx, N, kind):
    mesh = dgc_amd.Mesh.structured(14, 11, -1, 1, -1, 1, jitter=0.15, seed=N)
    disc = dgc_amd.Discretization(ctx, mesh, N)
    ip = dgc_amd.IPDG(disc, tau_mode=1)
    ip.set_boundary(kind)
    rp, ci, va = ip.assemble()
    n = disc.size
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    # row block c spans the cells within two faces: at most 10 cells
    assert (np.diff(rp) <= 10 * disc.Np).all()
    v = np.random.default_rng(N + 10 * kind).standard_normal(n)
    d = dgc_amd.DeviceArray.from_host(ctx, v)
    ip.apply(d)
    assert rel(A @ v, d.download()) < 1e-13
