"""lare_dev_step_info: a step with the per-step scalars and energies read
back in one copy, against the CPU restatement; and the state check that
refuses a step whose input holds a non-finite value or rho <= 0.

The faulty states are built on the host and uploaded: every kernel runs to
its normal end on them, and the library reports the fault from the
reduction that set_dt already makes (no device-side trap)."""
import numpy as np
import pytest

from lare3d_amd import configs
from lare3d_amd.abi import DeviceError

pytestmark = pytest.mark.gpu


def test_step_info_matches_restatement():
    from lare3d_amd.device import Device
    from oracle.oracle import Oracle
    grid, params, f = configs.harris(16)
    print("\n[case] step_info Harris 16^3, 4 steps", flush=True)
    dev = Device(grid, params, f)
    orc = Oracle(grid, params, f)
    try:
        for s in (dev, orc):
            s.boundary_conditions()
            s.eta_calc()
        for step in range(1, 5):
            info = dev.step_info(step, with_energy=True)
            dt, t = orc.step(step)
            en = orc.energy_account()
            so = orc.scalars()
            print(f"step {step}: dt {info.dt:.16e} substeps {info.substeps} en {info.en_b:.10e} "
                  f"{info.en_ke:.10e} {info.en_int:.10e} divb {info.max_divb:.2e}", flush=True)
            assert info.dt == dt and info.time == t and info.with_energy == 1
            assert info.substeps == so.last_substeps and info.dtr == so.dtr
            for a, b in zip((info.en_b, info.en_ke, info.en_int), en):
                assert abs(a - b) <= 1e-12 * abs(b)
            assert abs(info.total_visc_heating - so.total_visc_heating) <= 1e-10 * abs(so.total_visc_heating) + 1e-300
            assert abs(info.total_ohmic_heating - so.total_ohmic_heating) <= 1e-10 * abs(so.total_ohmic_heating)
            assert info.max_divb < 1e-12
        info = dev.step_info(5, with_energy=False)
        assert info.with_energy == 0 and info.en_b == 0.0
    finally:
        dev.close()
        orc.close()


@pytest.mark.parametrize("field,cell,value,text", [
    ("rho", (3, 4, 5), np.nan, "rho is not finite or not positive at (3, 4, 5)"),
    ("rho", (2, 2, 2), -1.0, "rho is not finite or not positive at (2, 2, 2)"),
    ("energy", (16, 1, 9), np.inf, "energy is not finite at (16, 1, 9)"),
    ("vy", (7, 8, 9), np.nan, "vy is not finite at (7, 8, 9)"),
    ("bz", (5, 6, 7), np.nan, "bz is not finite at (5, 6, 7)"),
])
def test_state_check_names_field_and_cell(field, cell, value, text):
    from lare3d_amd.device import Device
    grid, params, f = configs.harris(16)
    print(f"\n[case] state check {field}{cell} = {value}", flush=True)
    dev = Device(grid, params, f)
    try:
        dev.boundary_conditions()
        dev.eta_calc()
        dev.step(1)
        g = dev.fields()
        g.view(field, *[c for c in cell for _ in (0, 1)])[...] = value
        dev.upload(g)
        with pytest.raises(DeviceError) as e:
            dev.step_info(2)
        print(e.value, flush=True)
        assert "status 1:" in str(e.value) and text in str(e.value)
    finally:
        dev.close()
