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


def test_noalloc_rejects_sizes_before_any_device_call():
    """Sizes beyond the fixed scratch array are refused with LARE_E_ARG
    before the first HIP call (so this runs without a GPU)."""
    lib = abi.load_library()
    dummy = (C.c_double * 8)()
    fields = abi.LareFields()
    for name in ("rho", "energy", "vx", "vy", "vz", "bx", "by", "bz"):
        setattr(fields, name, C.cast(dummy, C.POINTER(C.c_double)))
    en = (C.c_double * 3)()
    for n, msg in (((4097, 8, 8), b"outside 4..4096"), ((512, 512, 512), b"exceed the scratch maximum"),
                   ((3, 8, 8), b"outside 4..4096")):
        grid = Grid(*n, ((0, 1),) * 3, (1,) * 6)
        g = grid_struct(grid)
        assert lib.lare_dev_energy_account_noalloc(C.byref(g), C.byref(fields), en, None) == -1002
        assert msg in lib.lare_dev_last_error(), lib.lare_dev_last_error()
    grid = Grid(8, 8, 8, ((0, 1),) * 3, (1,) * 6)
    g = grid_struct(grid)
    fields.bz = None
    assert lib.lare_dev_energy_account_noalloc(C.byref(g), C.byref(fields), en, None) == -1002
    assert b"missing field" in lib.lare_dev_last_error()
