"""Restart dumps of the device state, and restarting from them.

The reference writes restart snapshots with the SDF library in write_file
(diagnostics.F90:123-472) and reads them back in restart_data
(setup.F90:518-688).  The SDF library is not part of this build, so dumps
use this module's own format, with the same content as a reference restart
dump (restart_flag = ALL(dump_mask(1:8)), diagnostics.F90:175):

* header: code name, step, time, restart flag;
* constants dt, time_prev and visc_heating (diagnostics.F90:191-196),
  which restart_data reads into dt_from_restart, time_prev and
  total_visc_heating (setup.F90:602-609);
* the grid xb_global(0:n), yb_global, zb_global (diagnostics.F90:198-199);
* Rho, Energy on cells 1..n; Vx, Vy, Vz on nodes 0..n; Bx on faces
  0..nx x 1..ny x 1..nz, By and Bz likewise (cell_subarray, node_subarray
  and b?_subarray of mpi_routines.F90), FP64.

File layout (little-endian):
    8 bytes   magic b"LARE3DMP"
    uint32    format version (1)
    uint32    length L of the JSON header
    L bytes   JSON header: code, step, time, restart, n, constants, grid and
              a block table (id, name, stagger, dims, lower bounds, byte
              offset); every float is stored as float.hex() so it reads back
              to the same bits
    blocks    raw FP64 arrays in Fortran order, each at a 64-byte aligned
              offset from the start of the file

Restarting (restart_device) follows lare3d.f90:56-68: the fields are read
into arrays whose ghost layers are then filled by boundary_conditions, and
eta by eta_calc.  Two modes:

* "reference": restart = 1 and dt_from_restart = the dumped dt, so the
  first set_dt returns it unchanged (lagran.F90:1078-1086), while the edge
  viscosity of that first step sees dt = 0, as after a restart of the
  reference (dt is only set by set_dt);
* "continue": the dumped dt becomes the previous step's dt and set_dt runs
  normally, so the next step is the one an uninterrupted run takes (the
  ghosts and eta are functions of the dumped interior).
"""
import json
import os
import struct

import numpy as np

from .fields import Fields

MAGIC = b"LARE3DMP"
VERSION = 1
CODE_NAME = "Lare3d"

# (block id, field name, name, stagger, index range per axis relative to n)
_BLOCKS = (
    ("Rho", "rho", "Fluid/Rho", "cell_centre", ((1, 0), (1, 0), (1, 0))),
    ("Energy", "energy", "Fluid/Energy", "cell_centre", ((1, 0), (1, 0), (1, 0))),
    ("Vx", "vx", "Velocity/Vx", "vertex", ((0, 0), (0, 0), (0, 0))),
    ("Vy", "vy", "Velocity/Vy", "vertex", ((0, 0), (0, 0), (0, 0))),
    ("Vz", "vz", "Velocity/Vz", "vertex", ((0, 0), (0, 0), (0, 0))),
    ("Bx", "bx", "Magnetic_Field/Bx", "face_x", ((0, 0), (1, 0), (1, 0))),
    ("By", "by", "Magnetic_Field/By", "face_y", ((1, 0), (0, 0), (1, 0))),
    ("Bz", "bz", "Magnetic_Field/Bz", "face_z", ((1, 0), (1, 0), (0, 0))),
)


class DumpError(RuntimeError):
    pass


def _ranges(n, rng):
    """Inclusive Fortran index ranges (lo, hi) per axis; hi = n + rng[a][1]."""
    return [(rng[a][0], n[a] + rng[a][1]) for a in range(3)]


def _fhex(x):
    return float(x).hex()


def write_dump(path, fields, grid, step, time, dt, time_prev=0.0, visc_heating=0.0):
    """Write a restart dump of host fields (Fields with the reference's
    bounds) at `step` and `time`."""
    n = tuple(fields.n)
    if tuple(grid.n) != n:
        raise DumpError(f"grid {grid.n} and fields {n} differ")
    blocks, arrays = [], []
    offset = 0
    for bid, fname, name, stagger, rng in _BLOCKS:
        r = _ranges(n, rng)
        a = np.asfortranarray(fields.view(fname, r[0][0], r[0][1], r[1][0], r[1][1], r[2][0], r[2][1]),
                              dtype="<f8")
        blocks.append({"id": bid, "name": name, "stagger": stagger, "dims": list(a.shape),
                       "lower": [r[0][0], r[1][0], r[2][0]], "dtype": "<f8", "nbytes": int(a.nbytes)})
        arrays.append(a)
    header = {
        "code": CODE_NAME, "step": int(step), "time": _fhex(time), "restart": True, "n": list(n),
        "constants": {"dt": _fhex(dt), "time_prev": _fhex(time_prev), "visc_heating": _fhex(visc_heating)},
        # xb_global(0:n) per axis (diagnostics.F90:139-160)
        "grid": [[_fhex(x) for x in grid.xb[a][2:n[a] + 3]] for a in range(3)],
        "blocks": blocks,
    }
    # the data offsets depend on the header length, which depends on them:
    # fix the table with placeholder offsets of the final width first
    for b in blocks:
        b["offset"] = 10 ** 15
    base = 16 + len(json.dumps(header).encode())
    start = (base + 63) // 64 * 64
    offset = start
    for b in blocks:
        b["offset"] = offset
        offset = (offset + b["nbytes"] + 63) // 64 * 64
    text = json.dumps(header).encode()
    if 16 + len(text) > start:
        raise DumpError("header grew past the data start")
    tmp = path + ".tmp"
    with open(tmp, "wb") as fh:
        fh.write(MAGIC + struct.pack("<II", VERSION, len(text)) + text)
        for b, a in zip(blocks, arrays):
            fh.seek(b["offset"])
            fh.write(a.tobytes(order="F"))
    os.replace(tmp, path)


def read_dump(path):
    """Read a dump: returns a dict with step, time, dt, time_prev,
    visc_heating, n, grid (xb_global per axis) and fields (a Fields whose
    dumped ranges hold the data; ghost layers zero)."""
    with open(path, "rb") as fh:
        head = fh.read(16)
        if len(head) != 16 or head[:8] != MAGIC:
            raise DumpError(f"{path}: not a Lare3d dump")
        version, hlen = struct.unpack("<II", head[8:])
        if version != VERSION:
            raise DumpError(f"{path}: format version {version}, expected {VERSION}")
        h = json.loads(fh.read(hlen).decode())
        if h.get("code") != CODE_NAME:
            raise DumpError(f"{path}: written by {h.get('code')!r}, not {CODE_NAME}")
        if not h.get("restart"):
            raise DumpError(f"{path}: not a restart dump")
        n = tuple(h["n"])
        f = Fields(*n)
        byid = {b["id"]: b for b in h["blocks"]}
        for bid, fname, _, _, rng in _BLOCKS:
            if bid not in byid:
                raise DumpError(f"{path}: block {bid} missing")
            b = byid[bid]
            r = _ranges(n, rng)
            dims = [r[a][1] - r[a][0] + 1 for a in range(3)]
            if b["dims"] != dims or b["dtype"] != "<f8":
                raise DumpError(f"{path}: block {bid} has dims {b['dims']}, expected {dims}")
            fh.seek(b["offset"])
            raw = fh.read(b["nbytes"])
            a = np.frombuffer(raw, dtype="<f8").reshape(dims, order="F")
            f.view(fname, r[0][0], r[0][1], r[1][0], r[1][1], r[2][0], r[2][1])[...] = a
    c = h["constants"]
    return {"step": h["step"], "time": float.fromhex(h["time"]), "dt": float.fromhex(c["dt"]),
            "time_prev": float.fromhex(c["time_prev"]), "visc_heating": float.fromhex(c["visc_heating"]),
            "n": n, "grid": [np.array([float.fromhex(x) for x in ax]) for ax in h["grid"]], "fields": f}


def dump_device(path, dev, step, time_prev=0.0):
    """Download the device state and write a restart dump of it."""
    s = dev.scalars()
    write_dump(path, dev.fields(), dev.grid, step, s.time, s.dt, time_prev, s.total_visc_heating)


def restart_state(path, grid, params, mode="reference"):
    """The fields, parameters and scalars a restart from `path` starts with
    (lare3d.f90:56-68, setup.F90:518-688).  Returns (fields, params, d,
    scalars) where scalars are to be set after the boundary fill."""
    d = read_dump(path)
    if tuple(grid.n) != d["n"]:
        raise DumpError(f"dump grid {d['n']} differs from {grid.n}")
    for a in range(3):
        if not np.array_equal(grid.xb[a][2:grid.n[a] + 3], d["grid"][a]):
            raise DumpError(f"dump grid differs from the run's grid on axis {a}")
    if mode == "reference":
        params.restart = 1
        params.dt_from_restart = d["dt"]
        scal = {"time": d["time"], "total_visc_heating": d["visc_heating"]}
    elif mode == "continue":
        params.restart = 0
        scal = {"time": d["time"], "total_visc_heating": d["visc_heating"], "dt": d["dt"], "first_set_dt": 0}
    else:
        raise ValueError(f"unknown restart mode {mode!r}")
    return d["fields"], params, d, scal


def restart_sim(make, path, grid, params, mode="reference"):
    """Start a simulation (make(grid, params, fields) -> Device or Oracle)
    from a dump: upload, boundary_conditions, eta_calc, scalars.  Returns
    (sim, dumped step)."""
    fields, params, d, scal = restart_state(path, grid, params, mode)
    sim = make(grid, params, fields)
    sim.boundary_conditions()
    sim.eta_calc()
    sim.set_scalars(**scal)
    return sim, d["step"]


def restart_device(path, grid, params, mode="reference"):
    """restart_sim on the device."""
    from .device import Device
    return restart_sim(Device, path, grid, params, mode)
