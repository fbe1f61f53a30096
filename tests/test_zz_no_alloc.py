"""dgc_rk4_step_noalloc (work vectors in the library's static device array)
against dgc_rk4_step with caller-owned work vectors: bitwise equal results
on two calls in a row, the second on a smaller problem, and a clear error
above the size limit. Called through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import dgc_amd
from dgc_amd.core import _ptrs

pytestmark = pytest.mark.gpu


def _noalloc_step(op, state, t, dt):
    L = dgc_amd.lib()
    L.dgc_rk4_step_noalloc.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.c_size_t, C.c_double, C.c_double]
    L.dgc_rk4_step_noalloc.restype = C.c_int
    return L.dgc_rk4_step_noalloc(op._h, _ptrs([a.ptr for a in state]), op.disc.size, t, dt)


def _case(ctx, mesh, N, maxwell, seed):
    disc = dgc_amd.Discretization(ctx, mesh, N)
    op = dgc_amd.MaxwellTM(disc) if maxwell else dgc_amd.Advection(disc, 1.0, 0.5, dgc_amd.Advection.BC_ZERO)
    u0 = np.random.default_rng(seed).standard_normal(op.F * disc.size)
    return disc, op, u0


def test_noalloc_step_equals_step_with_work_vectors(ctx):
    cases = [
        (dgc_amd.Mesh.structured(16, 16, -1, 1, -1, 1, jitter=0.1, seed=0), 4, True),   # K = 512, N = 4
        (dgc_amd.Mesh.structured(6, 5, 0, 1, 0, 1, jitter=0.05, seed=1, periodic=True), 2, False),  # smaller
    ]
    for i, (mesh, N, maxwell) in enumerate(cases):
        print(f"no-alloc step case {i}: K={mesh.K} N={N} maxwell={maxwell}", flush=True)
        disc, op, u0 = _case(ctx, mesh, N, maxwell, i)
        dt = 0.25 * mesh.min_edge / (N + 1) ** 2
        a = [dgc_amd.DeviceArray.from_host(ctx, f) for f in np.split(u0, op.F)]
        b = [dgc_amd.DeviceArray.from_host(ctx, f) for f in np.split(u0, op.F)]
        work = op.new_work()
        for step in range(3):
            op.rk4_step(a, step * dt, dt, work)
            assert _noalloc_step(op, b, step * dt, dt) == 0, dgc_amd.lib().dgc_last_error()
        ctx.sync()
        got = np.concatenate([x.download() for x in b])
        want = np.concatenate([x.download() for x in a])
        np.testing.assert_array_equal(got, want)
        assert np.all(np.isfinite(got))


def test_noalloc_step_rejects_sizes_above_the_work_area(ctx):
    # F * K * Np = 3 * 2 * 160 * 160 * 6 = 921600 fits; 3 * 2 * 300^2 * 6 = 3.24M does not
    mesh = dgc_amd.Mesh.structured(300, 300, -1, 1, -1, 1)
    disc = dgc_amd.Discretization(ctx, mesh, 2)
    op = dgc_amd.MaxwellTM(disc)
    state = [dgc_amd.DeviceArray(ctx, disc.size) for _ in range(3)]
    rc = _noalloc_step(op, state, 0.0, 1e-4)
    assert rc == 1
    assert b"64 MiB" in dgc_amd.lib().dgc_last_error()
