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
