"""Device (HIP, through the C-ABI) versus the CPU restatement, FP64.

Bar: relative L2 <= 1e-10 per field per step (the FP64 parity bar), and,
since both codes perform the same operations in the same order, every
element of every field array (ghost layers included) equal after each full
step and after the boundary fills.
"""
import numpy as np
import pytest

from lare3d_amd import BC_PERIODIC, BC_USER, configs
from tests.util import STATE, field_errors, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _pair(cfg):
    from lare3d_amd.device import Device
    from oracle.oracle import Oracle
    grid, params, fields = cfg[:3]
    print(f"\n[case] n={grid.n} bcs={grid.bcs}", flush=True)
    return Device(grid, params, fields), Oracle(grid, params, fields)


def _init(dev, orc):
    for d in (dev, orc):
        d.boundary_conditions()
        d.eta_calc()


# n = 32: rows wider than two 128-byte lines, several waves per y-z plane
@pytest.mark.parametrize("n", [6, 32])
@pytest.mark.parametrize("bcs", [(BC_PERIODIC,) * 6, (BC_USER,) * 6,
                                 (BC_PERIODIC, BC_PERIODIC, BC_USER, BC_USER, BC_PERIODIC, BC_PERIODIC)])
def test_boundary_conditions_bit_exact(bcs, n):
    grid, params, f = configs.uniform(n, bcs=bcs)
    rng = np.random.default_rng(7)
    for name in STATE:
        a = getattr(f, name)
        a[...] = rng.standard_normal(a.shape)
    dev, orc = _pair((grid, params, f))
    dev.boundary_conditions()
    orc.boundary_conditions()
    fd, fo = dev.fields(), orc.fields()
    for name in STATE:
        assert np.array_equal(getattr(fd, name), getattr(fo, name)), name
    dev.close()


def test_conduct_heat_arcade():
    grid, params, f, dt = configs.arcade_conduction(16, 16, 32, s_target=9)
    dev, orc = _pair((grid, params, f))
    _init(dev, orc)
    for d in (dev, orc):
        d.set_scalars(dt=dt)
    for it in range(3):
        dev.conduct_heat()
        orc.conduct_heat()
        sd, so = dev.scalars(), orc.scalars()
        assert sd.n_s_stages == so.n_s_stages and sd.n_s_stages > 1
        e = rel_l2(dev.fields().energy, orc.fields().energy)
        print(f"super-step {it}: s={sd.n_s_stages} rel L2(energy incl. ghosts) = {e:.3e}")
        assert e <= TOL
    dev.close()


def _with(cfg, **kw):
    for k, v in kw.items():
        setattr(cfg[1], k, v)
    return cfg


CASES = {
    "orszag_tang": lambda: configs.orszag_tang(16),
    "harris": lambda: configs.harris(16),
    # non-uniform metrics on every axis (tanh stretch), so that a metric read
    # at a wrong index shows
    "harris_stretched": lambda: configs.harris(16, stretch=(48.0, 16.0, 48.0)),
    # viscous damping with a non-uniform visc3
    "harris_damped": lambda: configs.harris(16, damping=0.05),
    # odd, unequal sizes: partial tiles, waves and z-chunks on every axis
    "harris_odd": lambda: configs.harris(19, ny=13, nz=11),
    # a padded x extent of 69: two x tiles of the resistive kernel and two
    # waves per row of the x-sweep kernels, so neighbours cross tile and wave
    # edges inside the row
    "harris_wide": lambda: configs.harris(64, ny=9, nz=6),
    "arcade": lambda: configs.arcade_conduction(16, 16, 24, s_target=7),
    "arcade_odd": lambda: configs.arcade_conduction(13, 11, 19, s_target=5),
    # the twisted flux tube (walls in x and y, periodic z; resistive with
    # Braginskii conduction)
    "kink": lambda: configs.kink_tube(16),
    # the optional physics: the cooling_term (viscous and ohmic, with the
    # reference's stale-index average), the Boris correction (va_max below
    # the Alfven speed of the sheet) and the shock limiter; and on the
    # vortex the SHOCKCOMPRESSION variant with the cooling
    "harris_options": lambda: _with(configs.harris(16, ny=13, nz=11), cooling_term=1, alpha_av=0.05, boris=1,
                                    va_max=0.6, shock_limiter=1),
    "ot_compression": lambda: _with(configs.orszag_tang(16), shock_compression=1, cooling_term=1, alpha_av=0.3),
    # the remap kinetic-energy correction: odd sizes (chunks of the marching
    # momentum remap that end inside the edge range) and a wide x extent
    # (edge terms whose neighbours sit in the next wave)
    "harris_rke": lambda: _with(configs.harris(19, ny=13, nz=11), rke=1),
    "wide_rke": lambda: _with(configs.harris(64, ny=9, nz=6), rke=1),
    "ot_rke": lambda: _with(configs.orszag_tang(16), rke=1),
    # the CAUCHY predictor (cell-centred Cauchy solution for the half-step B)
    "harris_cauchy": lambda: _with(configs.harris(19, ny=13, nz=11), cauchy=1),
    "ot_cauchy": lambda: _with(configs.orszag_tang(16), cauchy=1, rke=1),
    # FOURTHORDER resistive substeps, alone and with conduction
    "harris_rk4": lambda: _with(configs.harris(19, ny=13, nz=11), fourth_order=1),
    "kink_rk4": lambda: _with(configs.kink_tube(16), fourth_order=1),
}


@pytest.mark.parametrize("case", list(CASES))
def test_full_steps(case):
    dev, orc = _pair(CASES[case]())
    _init(dev, orc)
    for step in range(1, 8):
        dtd, td = dev.step(step)
        dto, to = orc.step(step)
        fd, fo = dev.fields(), orc.fields()
        errs = field_errors(fd, fo)
        worst = max(errs.values())
        # whole arrays, ghost layers included, element for element
        ndiff = {n: int(np.count_nonzero(getattr(fd, n) != getattr(fo, n))) for n in STATE}
        print(f"{case} step {step}: dt dev {dtd:.16e} orc {dto:.16e} worst rel L2 {worst:.3e} "
              f"differing elements {sum(ndiff.values())}", flush=True)
        assert abs(dtd - dto) <= 1e-12 * abs(dto)
        assert worst <= TOL, errs
        assert sum(ndiff.values()) == 0, ndiff
    en_d = dev.energy_account()
    en_o = orc.energy_account()
    for a, b in zip(en_d, en_o):
        assert abs(a - b) <= 1e-12 * abs(b) + 1e-300
    sd, so = dev.scalars(), orc.scalars()
    assert abs(sd.total_visc_heating - so.total_visc_heating) <= 1e-10 * abs(so.total_visc_heating) + 1e-300
    assert abs(sd.total_ohmic_heating - so.total_ohmic_heating) <= 1e-10 * abs(so.total_ohmic_heating) + 1e-300
    assert sd.last_substeps == so.last_substeps and sd.n_s_stages == so.n_s_stages
    dev.close()


def test_harris_128_two_steps():
    """The bench workload's structure at 128^3 (many tiles per plane, many
    z-chunks and reduction blocks): two full steps, whole arrays equal to
    the oracle's element for element."""
    dev, orc = _pair(configs.harris(128))
    _init(dev, orc)
    for step in (1, 2):
        dtd, _ = dev.step(step)
        dto, _ = orc.step(step)
        fd, fo = dev.fields(), orc.fields()
        ndiff = {n: int(np.count_nonzero(getattr(fd, n) != getattr(fo, n))) for n in STATE}
        print(f"harris 128^3 step {step}: dt dev {dtd:.16e} orc {dto:.16e} differing elements {sum(ndiff.values())}",
              flush=True)
        assert dtd == dto
        assert sum(ndiff.values()) == 0, ndiff
    en_d, en_o = dev.energy_account(), orc.energy_account()
    for a, b in zip(en_d, en_o):
        assert abs(a - b) <= 1e-12 * abs(b)
    dev.close()
    orc.close()
