"""FP32 mode (LARE_PREC_FP32) against the FP64 mode of the same code.

FP32 mode stores the bulk fields in FP32 and computes the bulk arithmetic in
FP32 (Lagrangian step, resistive substep, the remap's geometry, mass, energy
and momentum fluxes, RKL2 stages).  FP64 stays for the face magnetic field
and its constrained-transport path (quotients, upwind values, fluxes and the
update), the Van Leer limiter's gradient comparisons, the metrics and all
reductions.  Bar (stated, not bit-exact): per-field relative L2 error against
the FP64 run after each step, and div B at FP64 round-off because the CT
update of B stays FP64.
"""
import numpy as np
import pytest

from lare3d_amd import configs
from lare3d_amd.abi import PREC_FP32, PREC_FP64
from tests.util import field_errors, rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "energy", "vx", "vy", "vz", "bx", "by", "bz")


def _run(make, precision, steps):
    from lare3d_amd.device import Device
    grid, params, f = make()
    params.precision = precision
    dev = Device(grid, params, f)
    out = []
    try:
        dev.boundary_conditions()
        dev.eta_calc()
        for s in range(1, steps + 1):
            dt, _ = dev.step(s)
            out.append((dt, dev.fields(), dev.energy_account(with_divb=True)))
    finally:
        dev.close()
    return out


@pytest.mark.parametrize("case,steps,tol", [("orszag_tang", 6, 2e-5), ("harris", 6, 2e-5)])
def test_fp32_mode_tracks_fp64(case, steps, tol):
    make = {"orszag_tang": lambda: configs.orszag_tang(32), "harris": lambda: configs.harris(32)}[case]
    print(f"\n[case] fp32 vs fp64 {case} 32^3, {steps} steps", flush=True)
    r64 = _run(make, PREC_FP64, steps)
    r32 = _run(make, PREC_FP32, steps)
    for s, ((dt64, f64, (en64, dvb64)), (dt32, f32, (en32, dvb32))) in enumerate(zip(r64, r32), 1):
        # scalars per field; vectors (v, B) as one field each, so that a
        # component that is near zero (bz of the Orszag-Tang vortex) is
        # measured against the magnitude of its vector
        errs = field_errors(f32, f64, ("rho", "energy"))
        for name, comps in (("v", ("vx", "vy", "vz")), ("B", ("bx", "by", "bz"))):
            a = np.concatenate([f32.interior(c).ravel() for c in comps])
            b = np.concatenate([f64.interior(c).ravel() for c in comps])
            errs[name] = rel_l2(a, b)
        worst = max(errs.values())
        print(f"step {s}: dt rel {abs(dt32 - dt64) / dt64:.2e}, worst field rel L2 {worst:.2e} "
              f"({max(errs, key=errs.get)}), div B {dvb32:.2e} (FP64 run {dvb64:.2e})")
        assert abs(dt32 - dt64) <= tol * dt64
        assert worst <= tol, errs
        for a, b in zip(en32, en64):
            assert abs(a - b) <= tol * abs(b) + 1e-12
        assert dvb32 <= 1e-12
