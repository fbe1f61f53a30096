"""Matrix-core operand layouts used by the MFMA kernels (bf16 32x32x8 and
32x32x16, fp32 16x16x4 and fp64 16x16x4) checked on the device against a host
product."""
import ctypes as C

import numpy as np
import pytest

import dgc_amd

pytestmark = pytest.mark.gpu


def _bf16_round(a):
    u = a.astype(np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_mfma_layout(ctx, kind):
    L = dgc_amd.lib()
    L.dgc_selftest_mfma.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(kind)
    m, k, n = {0: (32, 8, 32), 3: (32, 16, 32)}.get(kind, (16, 4, 16))
    A = rng.standard_normal((m, k)).astype(np.float32)
    B = rng.standard_normal((k, n)).astype(np.float32)
    Cm = np.zeros((m, n), np.float32)
    rc = L.dgc_selftest_mfma(ctx.handle, kind, A.ctypes.data, B.ctypes.data, Cm.ctypes.data)
    assert rc == 0, L.dgc_last_error()
    if kind in (0, 3):
        A, B = _bf16_round(A), _bf16_round(B)
    want = A.astype(np.float64) @ B.astype(np.float64)
    if kind == 2:
        # raw registers: find each (lane, register) value in the product and
        # check the documented fp64 layout, D register r of lane l = D[4r + l/16][l%16]
        raw = Cm.reshape(64, 4)
        found = {}
        for lane in range(64):
            for r in range(4):
                i, j = np.unravel_index(np.argmin(np.abs(want - raw[lane, r])), want.shape)
                assert abs(want[i, j] - raw[lane, r]) < 1e-5
                found[(lane, r)] = (int(i), int(j))
        print({k: v for k, v in found.items() if k[0] in (0, 1, 16, 17, 33, 63)})
        for (lane, r), (i, j) in found.items():
            assert (i, j) == (4 * r + lane // 16, lane % 16), (lane, r, i, j)
        return
    np.testing.assert_allclose(Cm, want, rtol=1e-5, atol=1e-5)
