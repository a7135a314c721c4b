"""The C-ABI library builds, loads, and exports every entry point declared
in include/lare3d_dev.h; argument and call-order errors are reported
without touching a GPU."""
import ctypes as C
import os
import re

import pytest

from lare3d_amd import abi
from lare3d_amd.configs import default_params
from lare3d_amd.grid import Grid, grid_struct

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "lare3d_dev.h")).read()
    return sorted(set(re.findall(r"\b(lare_dev_\w+)\s*\(", txt)))


def test_header_and_ctypes_agree():
    assert _declared() == sorted(abi.DEVICE_SYMBOLS)


def test_library_exports_all_symbols():
    lib = abi.load_library()
    for name in _declared():
        assert hasattr(lib, name), name


def test_call_order_and_argument_errors():
    lib = abi.load_library()
    fields = abi.LareFields()
    # upload before init
    assert lib.lare_dev_upload(C.byref(fields)) == -1001
    assert b"before lare_dev_init" in lib.lare_dev_last_error()
    # mixed boundary codes on one axis are rejected before any device call
    grid = Grid(8, 8, 8, ((0, 1),) * 3, (1, 2, 1, 1, 1, 1))
    g = grid_struct(grid)
    p = default_params()
    assert lib.lare_dev_init(C.byref(g), C.byref(p)) == -1002
    assert b"boundary codes" in lib.lare_dev_last_error()
    # FP32 bulk mode is a different entry configuration: unsupported here
    grid = Grid(8, 8, 8, ((0, 1),) * 3, (1,) * 6)
    g = grid_struct(grid)
    p = default_params(precision=7)
    assert lib.lare_dev_init(C.byref(g), C.byref(p)) == -1002
