"""Restart dumps (lare3d_amd/snapshot.py) on the CPU restatement.

* A dump reads back to the same bits: every dumped array and constant.
* "continue" restart: dumping after step 3 and restarting gives steps 4
  and 5 identical, element for element and ghost layers included, to the
  uninterrupted run -- the ghosts and eta are functions of the dumped
  interior, and the dumped dt is the previous dt the step needs.
* "reference" restart (setup.F90:518-688 with lagran.F90:1078-1086): the
  first set_dt returns the dumped dt without advancing the time, and that
  step deposits no viscous heat, because the edge viscosity sees dt = 0 as
  on step 1.
* Files that are not dumps, or belong to another grid, are refused.
"""
import os

import numpy as np
import pytest

from lare3d_amd import configs, snapshot
from lare3d_amd.fields import NAMES
from tests.backends import BACKENDS, close, make

CASES = {"harris": lambda: configs.harris(12, ny=10, nz=8), "kink": lambda: configs.kink_tube(8)}
COMPARE = ("rho", "energy", "vx", "vy", "vz", "bx", "by", "bz", "eta")


def _run(backend, cfg, steps, dump_after=None, path=None):
    grid, params, f = cfg
    sim = make(backend, grid, params, f)
    sim.boundary_conditions()
    sim.eta_calc()
    out = {}
    for s in range(1, steps + 1):
        out[s] = (sim.step(s), sim.fields())
        if s == dump_after:
            snapshot.write_dump(path, sim.fields(), grid, s, sim.scalars().time, sim.scalars().dt,
                                time_prev=0.25, visc_heating=sim.scalars().total_visc_heating)
    close(sim)
    return out


@pytest.mark.parametrize("case", list(CASES))
def test_dump_round_trip_bits(tmp_path, case):
    grid, params, f = CASES[case]()
    rng = np.random.default_rng(3)
    for name in NAMES:
        a = getattr(f, name)
        a[...] = rng.standard_normal(a.shape)
    p = str(tmp_path / "d.lare")
    snapshot.write_dump(p, f, grid, 17, 0.1 + 1e-17, 1.0 / 3.0, time_prev=0.2, visc_heating=np.pi)
    d = snapshot.read_dump(p)
    assert d["step"] == 17 and d["time"] == 0.1 + 1e-17 and d["dt"] == 1.0 / 3.0
    assert d["time_prev"] == 0.2 and d["visc_heating"] == np.pi
    nx, ny, nz = grid.n
    g = d["fields"]
    for name, r in (("rho", (1, nx, 1, ny, 1, nz)), ("energy", (1, nx, 1, ny, 1, nz)), ("vx", (0, nx, 0, ny, 0, nz)),
                    ("vz", (0, nx, 0, ny, 0, nz)), ("bx", (0, nx, 1, ny, 1, nz)), ("by", (1, nx, 0, ny, 1, nz)),
                    ("bz", (1, nx, 1, ny, 0, nz))):
        assert np.array_equal(g.view(name, *r), f.view(name, *r)), name
    for a in range(3):
        assert np.array_equal(d["grid"][a], grid.xb[a][2:grid.n[a] + 3])


def _restart_matches(backend, case, tmp_path):
    p = str(tmp_path / f"{backend}_{case}.lare")
    ref = _run(backend, CASES[case](), 5, dump_after=3, path=p)
    grid, params, _ = CASES[case]()
    sim, step = snapshot.restart_sim(lambda g_, p_, f_: make(backend, g_, p_, f_), p, grid, params, "continue")
    assert step == 3
    for s in (4, 5):
        dt, t = sim.step(s)
        (dt0, t0), f0 = ref[s]
        f1 = sim.fields()
        print(f"{case} step {s}: dt {dt:.17e} (uninterrupted {dt0:.17e})")
        assert dt == dt0 and t == t0
        for name in COMPARE:
            assert np.array_equal(getattr(f1, name), getattr(f0, name)), (s, name)
    close(sim)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("backend", BACKENDS)
def test_continue_restart_equals_uninterrupted(backend, case, tmp_path):
    _restart_matches(backend, case, tmp_path)


@pytest.mark.parametrize("backend", BACKENDS)
def test_reference_restart_semantics(backend, tmp_path):
    p = str(tmp_path / "h.lare")
    ref = _run(backend, CASES["harris"](), 3, dump_after=3, path=p)
    d = snapshot.read_dump(p)
    grid, params, _ = CASES["harris"]()
    sim, step = snapshot.restart_sim(lambda g_, p_, f_: make(backend, g_, p_, f_), p, grid, params, "reference")
    dt, t = sim.step(step + 1)
    # set_dt returns before time = time + dt (lagran.F90:1083-1084, :1164):
    # the first step after a restart does not advance the time
    assert dt == d["dt"] and t == d["time"]
    # dt = 0 before the first set_dt: no viscous heating on that step
    assert not np.any(sim.fields().visc_dep)
    assert sim.scalars().total_visc_heating == d["visc_heating"]
    dt2, _ = sim.step(step + 2)
    assert dt2 != dt and np.isfinite(dt2)
    assert np.any(sim.fields().visc_dep)
    close(sim)


def test_refuses_foreign_files(tmp_path):
    p = str(tmp_path / "x.lare")
    with open(p, "wb") as fh:
        fh.write(b"SDF1" + bytes(60))
    with pytest.raises(snapshot.DumpError, match="not a Lare3d dump"):
        snapshot.read_dump(p)
    grid, params, f = configs.harris(12, ny=10, nz=8)
    snapshot.write_dump(p, f, grid, 1, 0.0, 0.1)
    g2, p2, _ = configs.harris(12, ny=10, nz=9)
    with pytest.raises(snapshot.DumpError, match="differs"):
        snapshot.restart_state(p, g2, p2)
    g3, p3, _ = configs.harris(12, ny=10, nz=8, stretch=(30.0, None, None))
    with pytest.raises(snapshot.DumpError, match="grid differs"):
        snapshot.restart_state(p, g3, p3)
