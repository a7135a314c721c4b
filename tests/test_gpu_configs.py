"""The FP32 mode (FP32 bulk storage and arithmetic) on the conduction and
flux-tube configurations, against the FP64 mode of the same device code.

* Braginskii conduction on the sheared arcade (RKL2 super-steps with a
  fixed dt, FP32 bulk with FP64 reductions): per super-step relative L2 of
  the temperature, 20 super-steps; and the surrogate's trapezoid residual of
  the plain RKL2 super-step in both precisions.
* The kink-unstable twisted flux tube (resistive MHD with Braginskii
  conduction, walls in x and y, periodic z): per-field relative L2 after
  each step, the energy budget (magnetic, kinetic, internal, and the total
  with the viscous and ohmic heating) and div B, which the FP64 CT update
  keeps at round-off in both modes.

Bars are stated per quantity; the FP32 mode is not bit-exact by design.
"""
import numpy as np
import pytest

from lare3d_amd import configs
from lare3d_amd.abi import PREC_FP32, PREC_FP64
from tests.util import rel_l2

pytestmark = pytest.mark.gpu


def _conduct(precision, n, supersteps):
    from lare3d_amd.device import Device
    grid, params, f, dt = configs.arcade_conduction(*n, s_target=15)
    params.precision = precision
    dev = Device(grid, params, f)
    out = []
    try:
        dev.boundary_conditions()
        dev.set_scalars(dt=dt)
        for _ in range(supersteps):
            dev.conduct_heat()
            out.append((dev.scalars().n_s_stages, dev.fields().interior("energy").copy()))
    finally:
        dev.close()
    return out, f.interior("energy").copy()


def test_arcade_conduction_fp32_vs_fp64():
    n = (32, 32, 64)
    print(f"\n[case] arcade {n}, 20 RKL2 super-steps, FP32 vs FP64", flush=True)
    r64, e0 = _conduct(PREC_FP64, n, 20)
    r32, _ = _conduct(PREC_FP32, n, 20)
    for q, ((s64, e64), (s32, e32)) in enumerate(zip(r64, r32), 1):
        err = rel_l2(e32, e64)
        derr = rel_l2(e32 - e0, e64 - e0)
        print(f"super-step {q}: stages {s64}/{s32}, T rel L2 {err:.2e}, dT rel L2 {derr:.2e}", flush=True)
        assert s32 == s64 and s64 >= 3
        if q == 1:
            assert s64 >= 9
        assert err <= 1e-6
        assert derr <= 1e-4


def _kink(precision, n, steps):
    from lare3d_amd.device import Device
    grid, params, f = configs.kink_tube(*n)
    params.precision = precision
    dev = Device(grid, params, f)
    out = []
    try:
        dev.boundary_conditions()
        dev.eta_calc()
        for s in range(1, steps + 1):
            dt, _ = dev.step(s)
            en, divb = dev.energy_account(with_divb=True)
            sc = dev.scalars()
            out.append((dt, dev.fields(), en, divb, sc.total_visc_heating, sc.total_ohmic_heating,
                        sc.n_s_stages))
    finally:
        dev.close()
    return out


def test_kink_tube_fp32_vs_fp64():
    n = (32, 32, 64)
    steps = 20
    print(f"\n[case] kink tube {n}, {steps} steps, FP32 vs FP64", flush=True)
    r64 = _kink(PREC_FP64, n, steps)
    r32 = _kink(PREC_FP32, n, steps)
    e_start = None
    for s, (a, b) in enumerate(zip(r64, r32), 1):
        dt64, f64, en64, dvb64, vh64, oh64, st64 = a
        dt32, f32, en32, dvb32, vh32, oh32, st32 = b
        errs = {nm: rel_l2(f32.interior(nm), f64.interior(nm)) for nm in ("rho", "energy")}
        for nm, comps in (("v", ("vx", "vy", "vz")), ("B", ("bx", "by", "bz"))):
            errs[nm] = rel_l2(np.concatenate([f32.interior(c).ravel() for c in comps]),
                              np.concatenate([f64.interior(c).ravel() for c in comps]))
        tot64, tot32 = sum(en64), sum(en32)
        if e_start is None:
            e_start = tot64
        print(f"step {s}: dt rel {abs(dt32 - dt64) / dt64:.1e} stages {st64}/{st32} "
              + " ".join(f"{k} {v:.1e}" for k, v in errs.items())
              + f" | E_b {en64[0]:.6e} E_k {en64[1]:.4e} E_i {en64[2]:.6e}"
              f" total drift {(tot64 - e_start) / e_start:+.2e} (FP32 {(tot32 - e_start) / e_start:+.2e})"
              f" | heating visc {vh64:.3e} ohmic {oh64:.3e} | div B {dvb64:.1e}/{dvb32:.1e}", flush=True)
        assert abs(dt32 - dt64) <= 1e-5 * dt64
        assert errs["rho"] <= 1e-5 and errs["energy"] <= 1e-5 and errs["B"] <= 1e-5
        assert errs["v"] <= 1e-3
        for x, y in zip(en32, en64):
            assert abs(x - y) <= 1e-5 * abs(y) + 1e-7 * tot64
        assert abs(tot64 - e_start) <= 1e-3 * e_start
        assert dvb64 <= 1e-12 and dvb32 <= 1e-12


def _plain_residual(precision):
    from lare3d_amd.device import Device
    from lare3d_amd.surrogate import Surrogate
    grid, params, f, dt = configs.harris_conduction(32, seed=999, s_target=5)
    params.precision = precision
    dev = Device(grid, params, f)
    try:
        dev.boundary_conditions()
        dev.set_scalars(dt=dt)
        return Surrogate(dev).plain_residual()
    finally:
        dev.close()


def test_trapezoid_residual_fp32_vs_fp64():
    """The residual check of the conduction surrogate evaluates the
    trapezoid with the stage kernel (all history weights zero).  Applied to
    the plain RKL2 super-step it must read the same in both precisions."""
    print("\n[case] trapezoid residual of plain RKL2, Harris + conduction 32^3, FP32 vs FP64", flush=True)
    r64 = _plain_residual(PREC_FP64)
    r32 = _plain_residual(PREC_FP32)
    print(f"residual FP64 {r64:.4e}, FP32 {r32:.4e}")
    assert 0.0 < r64 < 1e-4
    assert abs(r32 - r64) <= 0.05 * r64 + 1e-6
