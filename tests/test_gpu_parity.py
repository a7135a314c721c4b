"""Device (HIP, through the C-ABI) versus the CPU restatement, FP64.

Bar: relative L2 <= 1e-10 per field per step (the FP64 parity bar), and,
since both codes perform the same operations in the same order, every
element of every field array (ghost layers included) equal after each full
step and after the boundary fills.
"""
import numpy as np
import pytest

from lare3d_amd import BC_OPEN, BC_PERIODIC, BC_USER, configs
from lare3d_amd.abi import EOS_ION, EOS_PI
from tests.util import STATE, field_errors, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _pair(cfg):
    from lare3d_amd.device import Device
    from oracle.oracle import Oracle
    grid, params, fields = cfg[:3]
    print(f"\n[case] n={grid.n} bcs={grid.bcs}", flush=True)
    pair = Device(grid, params, fields), Oracle(grid, params, fields)
    for d in pair:
        d.pi = params.eos_number in (EOS_ION, EOS_PI)
        d.cowling = bool(params.cowling_resistivity)
    return pair


def _init(dev, orc):
    """lare3d.f90:65-74: boundary_conditions, eta_calc, and for a partially
    ionised run neutral_fraction and perpendicular_resistivity."""
    for d in (dev, orc):
        d.boundary_conditions()
        d.eta_calc()
        if d.pi:
            d.neutral_fraction()
        if d.cowling:
            d.perpendicular_resistivity()


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


@pytest.mark.parametrize("eos", [EOS_PI, EOS_ION])
def test_neutral_fraction_bit_exact(eos):
    """neutral_fraction over the whole cell box and perpendicular_resistivity
    at nodes 0..n on a random state across the ionisation transition
    (energy and rho random per cell, B random per face)."""
    grid, params, f = configs.partially_ionised(configs.uniform(8, bcs=(BC_USER,) * 6), eos, 1)
    rng = np.random.default_rng(11 + eos)
    f.rho[...] = 10.0 ** rng.uniform(-1.0, 1.0, f.rho.shape)
    f.energy[...] = rng.uniform(0.2, 6.0, f.energy.shape)
    for name in ("bx", "by", "bz"):
        getattr(f, name)[...] = rng.standard_normal(getattr(f, name).shape)
    dev, orc = _pair((grid, params, f))
    for d in (dev, orc):
        d.neutral_fraction()
        d.perpendicular_resistivity()
    (xd, ed), (xo, eo) = dev.neutral(True, True), orc.neutral(True, True)
    print(f"eos {eos}: xi {xo.min():.3f}..{xo.max():.3f}, eta_perp {eo.max():.3e}; differing xi "
          f"{int(np.count_nonzero(xd != xo))}, eta_perp {int(np.count_nonzero(ed != eo))}", flush=True)
    assert np.array_equal(xd, xo)
    assert np.array_equal(ed, eo)
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


def _open(cfg, bcs):
    from lare3d_amd.grid import Grid
    g = cfg[0]
    grid = Grid(*g.n, g.extent, bcs, g.stretch, g.widths)
    return (grid,) + tuple(cfg[1:])


def _alfven_box(n=(14, 11, 9)):
    """x and y sides open, z periodic; B along x plus a transverse field and
    velocity noise confined to the middle of the box, so that the x far
    field has no perpendicular field; resistive, conducting, rke."""
    from lare3d_amd.fields import Fields
    from lare3d_amd.grid import Grid
    grid = Grid(*n, ((0.0, 1.0), (0.0, 0.8), (0.0, 0.6)), (BC_OPEN, BC_OPEN, BC_OPEN, BC_OPEN, BC_PERIODIC,
                                                          BC_PERIODIC))
    params = configs.default_params(resistive_mhd=1, eta_background=2e-3, eta0=0.0, j_max=1e30, conduction=1,
                                    kappa_0=2e-2, heat_flux_limiter=1, rke=1)
    f = Fields(*n)
    rng = np.random.default_rng(21)
    f.rho[...] = 1.0
    f.energy[...] = 0.6 / (params.gamma - 1.0)
    f.bx[...] = 0.8
    mid = lambda a, lo: np.abs(np.arange(a.shape[0])[:, None, None] + lo - n[0] / 2) < n[0] / 4
    for name, lo in (("by", -1), ("bz", -1), ("vy", -2), ("vz", -2), ("vx", -2)):
        a = getattr(f, name)
        a[...] = np.where(mid(a, lo), 0.05 * rng.standard_normal(a.shape), 0.0)
    return grid, params, f


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
    # the CT sweep's tiles: 58 + 5 = 2 x 31 + 1 points across x in the y and
    # z sweeps, 10 + 5 = 2 x 7 + 1 rows and 12 + 5 = 2 x 8 + 1 points along
    # the march (a last tile or chunk of one point); 56 + 5 = 60 + 1 points
    # across x in the x sweep
    "harris_tile_edges": lambda: configs.harris(58, ny=10, nz=12),
    "harris_tile_edges_x": lambda: configs.harris(56, ny=16, nz=20),
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
    # the partially ionised equation of state (neutral fraction per cell)
    # with Cowling resistivity: across the transition on the sheet; with the
    # ionisation energy in the pressure and conduction (nine RKL2 stages,
    # the neutral fraction recomputed between them); on the flux tube with
    # FOURTHORDER substeps; and the ionisation energy on the ideal vortex,
    # through the predictor remap and through the CAUCHY predictor
    "harris_pi_cowling": lambda: configs.partially_ionised(configs.harris(19, ny=13, nz=11), EOS_PI, 1),
    "harris_ion_cond": lambda: configs.partially_ionised(
        _with(configs.harris(16), conduction=1, kappa_0=100.0, heat_flux_limiter=1), EOS_ION, 1),
    "kink_pi_rk4": lambda: configs.partially_ionised(_with(configs.kink_tube(16), fourth_order=1), EOS_PI, 1,
                                                     temp_norm=4.0e5, L_norm=1.0e3),
    "ot_ion": lambda: configs.partially_ionised(configs.orszag_tang(16), EOS_ION, 0, temp_norm=2.0e5),
    "ot_ion_cauchy": lambda: configs.partially_ionised(_with(configs.orszag_tang(16), cauchy=1), EOS_ION, 0,
                                                       temp_norm=2.0e5),
    # open (far-field characteristic) boundaries: the sheet's y walls open
    # (B along x, so the fast solver); the flux tube open in x and y (four
    # open sides meeting at edges) with conduction; a box open in x with a
    # field along x and no perpendicular far field (the Alfven solver there,
    # the fast one on its open y sides), resistive with conduction and rke
    "harris_open": lambda: _open(configs.harris(19, ny=13, nz=11), (BC_PERIODIC, BC_PERIODIC, BC_OPEN, BC_OPEN,
                                                                    BC_PERIODIC, BC_PERIODIC)),
    "kink_open": lambda: _open(configs.kink_tube(16), (BC_OPEN, BC_OPEN, BC_OPEN, BC_OPEN, BC_PERIODIC,
                                                       BC_PERIODIC)),
    "alfven_open": lambda: _alfven_box(),
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
        if dev.pi:
            # neutral fraction and Cowling resistivity, element for element
            (xd, ed), (xo, eo) = dev.neutral(True, dev.cowling), orc.neutral(True, dev.cowling)
            assert np.array_equal(xd, xo), np.abs(xd - xo).max()
            if dev.cowling:
                assert np.array_equal(ed, eo), np.abs(ed - eo).max()
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
