"""lare_dev_energy_account_noalloc (the entry point that allocates nothing)
against lare_dev_energy_account on the same field values.

The fields are first advanced by the library (so that v, B and the ghost
layers are those of a real step), downloaded, and copied into device
buffers that the caller owns (torch, as device-memory plumbing only) in the
Fortran layout of the C-ABI.  The buffers are allocated through the HIP
runtime the library itself links (a second runtime in the process, such
as torch's own copy, would not see the device).  The no-allocation entry point then reads those
buffers; its kernels are the library's energy kernels, so the energies and
the div B diagnostic must be the same bits.  Two calls in a row, the second
on a smaller grid, check that nothing of the first call's scratch contents
leaks into the second.
"""
import ctypes as C

import numpy as np
import pytest

from lare3d_amd import configs
from lare3d_amd.abi import LareFields, check, load_library
from lare3d_amd.grid import grid_struct

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)
STATE = ("rho", "energy", "vx", "vy", "vz", "bx", "by", "bz")


def _library_energies(cfg, steps):
    from lare3d_amd.device import Device
    grid, params, f = cfg
    dev = Device(grid, params, f)
    try:
        dev.boundary_conditions()
        dev.eta_calc()
        for s in range(1, steps + 1):
            dev.step(s)
        en, divb = dev.energy_account(with_divb=True)
        fields = dev.fields()
    finally:
        dev.close()
    return grid, fields, en, divb


def _hip():
    """The HIP runtime that the device library links (already loaded)."""
    load_library()
    with open("/proc/self/maps") as fh:
        path = next(ln.split()[-1] for ln in fh if "libamdhip64.so" in ln)
    hip = C.CDLL(path)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _noalloc(grid, fields):
    lib = load_library()
    hip = _hip()
    bufs = {}
    try:
        # caller-owned device copies of the Fortran arrays, same memory order
        for n in STATE:
            a = getattr(fields, n)
            assert a.flags.f_contiguous and a.dtype == np.float64
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), a.nbytes) == 0
            bufs[n] = p
            assert hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        st = LareFields()
        for n in STATE:
            setattr(st, n, C.cast(bufs[n], _dp))
        g = grid_struct(grid)
        en = (C.c_double * 3)()
        dvb = C.c_double()
        check(lib, lib.lare_dev_energy_account_noalloc(C.byref(g), C.byref(st), en, C.byref(dvb)),
              "energy_account_noalloc")
    finally:
        for p in bufs.values():
            hip.hipFree(p)
    return tuple(en), dvb.value


def test_noalloc_energy_matches_library_twice():
    cases = [("Harris 24x20x16, 3 steps", lambda: configs.harris(24, ny=20, nz=16), 3),
             ("Orszag-Tang 12^3, 2 steps", lambda: configs.orszag_tang(12), 2)]
    for label, make, steps in cases:
        print(f"\n[case] {label}", flush=True)
        grid, fields, en_ref, divb_ref = _library_energies(make(), steps)
        en, divb = _noalloc(grid, fields)
        print(f"library {en_ref} {divb_ref:.3e}\nnoalloc {en} {divb:.3e}", flush=True)
        assert en == en_ref
        assert divb == divb_ref
        assert en_ref[0] > 0 and en_ref[1] > 0 and en_ref[2] > 0
