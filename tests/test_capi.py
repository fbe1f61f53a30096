"""The C-ABI library loads and exports every entry point include/*.h
declares; argument checking that needs no GPU."""
import ctypes
import glob
import os
import re

import pytest

import dgc_amd
from conftest import ROOT


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(dgc_[A-Za-z0-9_]+)\s*\(",
                             text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared_functions()
    assert "dgc_rk4_step" in names and "dgc_disc2d_create" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(dgc_amd.lib_path())
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_adv1d_operators_on_the_host():
    import numpy as np

    D, L, x = dgc_amd.Advection1D.operators(4, 2)
    # Dr annihilates constants, V V^T is symmetric, the nodes span [0, 2 pi]
    assert np.abs(D.sum(axis=1)).max() < 1e-12
    assert np.abs(L - L.T).max() < 1e-12
    assert abs(x.min()) < 1e-15 and abs(x.max() - 2 * np.pi) < 1e-14


def test_max_order():
    assert dgc_amd.lib().dgc_max_order() == 7


def test_null_and_range_checks_without_gpu():
    L = dgc_amd.lib()
    assert L.dgc_ctx_create(0, None) == 1
    assert b"null" in L.dgc_last_error()
    assert L.dgc_op_rhs(None, None, 0, 0.0, None) == 1
    assert L.dgc_rk4_step(None, None, 0, 0.0, 0.0, None) == 1
    assert L.dgc_disc2d_create(None, 0, 1, *([None] * 8), 0, None, None, None, None, None) == 1
    assert L.dgc_ref_operators(9, 0, *([None] * 7)) == 0  # any order >= 1 builds on the host
    assert L.dgc_ref_operators(1, 5, *([None] * 7)) == 1
    assert L.dgc_ctx_sync(None) == 1
    assert L.dgc_axpy(None, 1.0, None, 0, None, None) == 1
    assert L.dgc_disc2d_create_from_cells(None, 8, 9, 1, None, None, None, None) == 1
    assert b"null" in L.dgc_last_error()
    assert L.dgc_disc2d_arrays(None, None, None, None, None, None, None, None) == 1
    # widened rows: every entry point refuses a null handle before touching a device
    assert L.dgc_graddiv_create(None, None, None, None, None) == 1
    assert L.dgc_graddiv_set_tiling(None, 1) == 1
    assert L.dgc_grad2d_apply(None, None, None, 1.0, None, None) == 1
    assert L.dgc_div2d_apply(None, None, None, None, None, 1.0, None) == 1
    assert L.dgc_ipdg_set_boundary(None, 1) == 1
    assert L.dgc_ipdg_apply_inhomogeneous(None, None, None, None) == 1
    assert L.dgc_ipdg_apply_data(None, None, None, None) == 1
    assert L.dgc_ipdg_assemble(None, None, None, None, None) == 1
    assert L.dgc_ipdg_set_shift(None, 1.0) == 1
    assert L.dgc_stokes_create(None, None) == 1
    assert L.dgc_stokes_coarse_space(None, None, None, 2, 2) == 1
    assert L.dgc_stokes_advance(None, None, None, 0.0, None, None, 0.0, 1e-3, 1e-10, None, None, None) == 1
    assert L.dgc_adv1d_create(None, 8, 2, None, None, None) == 1
    assert L.dgc_adv1d_rhs(None, None, 0.0, None) == 1
    assert L.dgc_adv1d_rk4_steps(None, None, 0.0, 1e-3, 1) == 1
    assert L.dgc_adv1d_operators(8, 13, None, None, None) == 1  # order beyond 12
    assert L.dgc_write_plot_file(None, 0, None, None, None, None) == 1
