"""The Fortran driver (lare3d_amd/fortran/ot_driver.f90) calls the device
library through the ISO_C_BINDING bridge; its dt, time and energies match
the CPU restatement run from the same initial condition."""
import os
import subprocess

import numpy as np
import pytest

from lare3d_amd import configs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "lare3d_amd", "fortran", "ot_driver")


def test_fortran_driver_matches_restatement():
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (amdflang absent at build time)")
    from oracle.oracle import Oracle
    print("\n[case] Fortran ot_driver 16 5", flush=True)
    out = subprocess.run([DRIVER, "16", "5"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    rows = [l.split() for l in out.stdout.splitlines() if l.startswith("step")]
    assert len(rows) == 5
    grid, params, fields = configs.orszag_tang(16)
    orc = Oracle(grid, params, fields)
    orc.boundary_conditions()
    orc.eta_calc()
    for r in rows:
        step = int(r[1])
        dt, t = orc.step(step)
        en = orc.energy_account()
        vals = [float(x) for x in r[2:]]
        print(f"step {step}: fortran dt {vals[0]:.15e} oracle {dt:.15e}")
        assert abs(vals[0] - dt) <= 1e-12 * dt and abs(vals[1] - t) <= 1e-12 * t
        for a, b in zip(vals[2:], en):
            assert abs(a - b) <= 1e-10 * abs(b)
