"""Pins of the optional physics of lagrangian_step: the SHOCKLIMITER and
SHOCKCOMPRESSION variants of edge_viscosity (lagran.F90:801-847), the Boris
correction (:1092-1108, :375-377, :556) and the cooling_term of the viscous
heating (:464-476).  Each is a closed form that the reference's
discretisation fixes, checked on the oracle and (gpu) on the device.

* Linear compression vx = -a x: every x edge has the same dv, so the
  limiter's ratios are rl = rr = 1 and psi = 1: the viscosity vanishes and
  set_dt takes the inviscid step.  Without the limiter it takes the
  Kuropatenko step (test_pin_oracle3d).
* Linear expansion vx = +a x: the reference's default viscosity acts on
  expanding edges too (dvdots = -|dvdots|), SHOCKCOMPRESSION switches it off
  (dvdots = MIN(0, dvdots)); on compression both agree.
* Boris: on a uniform state the CFL step is
  0.9 min(dx / sqrt((gamma p + gamma_b B^2) / rho), dx / |v|) with
  gamma_b = 1 / (1 + B^2 / (rho va_max^2)) once B^2 / rho >= va_max^2,
  and a huge va_max gives the step without it, bit for bit.
* cooling_term: the first cooled step removes
  alpha_av dt visc_heat / rho_new = alpha_av (visc_dep increment) / rho_new
  from every cell, and nothing else changes.
"""
import math

import numpy as np
import pytest

from lare3d_amd import BC_PERIODIC, BC_USER, configs
from lare3d_amd.configs import default_params
from lare3d_amd.fields import Fields
from lare3d_amd.grid import Grid
from tests.backends import BACKENDS, close, make
from tests.physics import along

GAMMA = 5.0 / 3.0


def _line_dt(backend, sign, quad=0.0, **variant):
    """dt of set_dt after one lagrangian_step of vx = sign * a * x (- quad x^2)
    with the previous dt well above the switch-on threshold; also the
    inviscid and, for quad = 0, the Kuropatenko dt of the setup (as in
    test_edge_viscosity_kuropatenko_dt)."""
    n, a, p0 = 16, 0.5, 0.6
    bvec = (0.3, 0.4, 0.2)
    grid = Grid(n, 4, 4, ((-0.5, 0.5), (0.0, 0.0625), (0.0, 0.0625)),
                (BC_USER, BC_USER, BC_PERIODIC, BC_PERIODIC, BC_PERIODIC, BC_PERIODIC))
    params = default_params(visc1=0.13, visc2=0.7, **variant)
    f = Fields(n, 4, 4)
    xc = grid.xc[0]
    rho_c = 1.0 + 0.4 * np.sin(2.0 * math.pi * xc)
    f.rho[...] = along(f.rho.shape, 0, rho_c)
    f.energy[...] = along(f.energy.shape, 0, p0 / ((GAMMA - 1.0) * rho_c))
    xb = grid.xb[0]
    vnode = sign * a * xb - quad * xb * xb
    f.vx[...] = along(f.vx.shape, 0, vnode)
    f.bx[...], f.by[...], f.bz[...] = bvec
    sim = make(backend, grid, params, f)
    sim.set_scalars(dt=1.0e-3)
    sim.lagrangian_step()
    dt = sim.scalars().dt
    vd = float(np.abs(sim.fields().visc_dep).max())
    close(sim)
    dx, length = 1.0 / n, 0.25 / n
    b2 = sum(b * b for b in bvec)
    cs = np.sqrt((GAMMA * p0 + b2) / rho_c)
    rho_v = 0.5 * (rho_c[:-1] + rho_c[1:])
    cs_v = 0.5 * (cs[:-1] + cs[1:])
    c1, c2 = params.visc1, 0.25 * (GAMMA + 1.0) * params.visc2
    best = nov = math.inf
    for i in range(0, n + 1):
        r1, r2 = rho_v[i], rho_v[i + 1]
        rho_e = 2.0 * r1 * r2 / (r1 + r2)
        c_e = min(cs_v[i], cs_v[i + 1])
        dv = a * dx
        q = rho_e * (c2 * dv + math.sqrt((c2 * dv) ** 2 + (c1 * c_e) ** 2)) * dv
        cv2 = q / rho_c[i + 1]
        dt1 = length / (math.sqrt(cv2) + math.sqrt(cs[i + 1] ** 2 + cv2))
        dt2 = dx / max(abs(vnode[i + 2]), 1e-300)
        best = min(best, dt1, dt2)
        nov = min(nov, length / cs[i + 1], dt2)
    return dt, vd, 0.9 * nov, 0.9 * best


@pytest.mark.parametrize("case,sign,variant,viscous", [
    ("default, expansion", 1.0, {}, True),
    ("SHOCKCOMPRESSION, expansion", 1.0, {"shock_compression": 1}, False),
    ("SHOCKCOMPRESSION, compression", -1.0, {"shock_compression": 1}, True),
    ("SHOCKLIMITER, linear compression", -1.0, {"shock_limiter": 1}, False),
])
@pytest.mark.parametrize("backend", BACKENDS)
def test_shock_variants(backend, case, sign, variant, viscous):
    dt, vd, inviscid, kuro = _line_dt(backend, sign, **variant)
    want = kuro if viscous else inviscid
    print(f"{case}: dt {dt:.15e} expected {want:.15e} (inviscid {inviscid:.6e}, viscous {kuro:.6e}); "
          f"max visc_dep {vd:.3e}")
    assert abs(dt - want) < 1e-12 * want
    assert inviscid - kuro > 1e-3 * kuro
    if not viscous:
        assert vd == 0.0


@pytest.mark.parametrize("backend", BACKENDS)
def test_shock_limiter_smooth_compression(backend):
    """vx = -a x - b x^2 (a = 0.5, b = 0.2): the edge dv grows linearly along
    x, so rl = 1 + d and rr = 1 - d with d < 1/2, 0.5 (rl + rr) = 1 up to
    rounding and psi = 1: the limited viscosity vanishes to round-off and
    set_dt takes the inviscid step.  A limiter that used one side twice
    (psi = min(2 rl, ...) or rr for rl) would leave a viscosity of order d."""
    dt, vd, inviscid, kuro = _line_dt(backend, -1.0, quad=0.2, shock_limiter=1)
    print(f"smooth compression with the limiter: dt {dt:.15e}, inviscid {inviscid:.15e}, max visc_dep {vd:.2e}")
    assert abs(dt - inviscid) < 1e-9 * inviscid
    assert vd < 1e-12


def test_shock_limiter_keeps_a_jump():
    """At a velocity jump the neighbouring edges carry no dv, so rl = rr = 0,
    psi = 0 and the limited viscosity equals the unlimited one: the viscous
    heating is identical, bit for bit."""
    out = []
    for lim in (0, 1):
        n = 16
        grid = Grid(n, 4, 4, ((-0.5, 0.5), (0.0, 0.0625), (0.0, 0.0625)),
                    (BC_USER, BC_USER, BC_PERIODIC, BC_PERIODIC, BC_PERIODIC, BC_PERIODIC))
        params = default_params(visc1=0.13, visc2=0.7, shock_limiter=lim)
        f = Fields(n, 4, 4)
        f.rho[...] = 1.0
        f.energy[...] = 0.9
        xb = grid.xb[0]
        f.vx[...] = along(f.vx.shape, 0, np.where(xb < 0.0, 0.3, -0.3))
        sim = make("oracle", grid, params, f)
        sim.set_scalars(dt=1.0e-3)
        sim.lagrangian_step()
        out.append((sim.scalars().dt, sim.fields()))
        close(sim)
    assert out[0][0] == out[1][0]
    assert np.any(out[0][1].visc_dep)
    for name in ("visc_dep", "energy", "vx", "rho"):
        assert np.array_equal(getattr(out[0][1], name), getattr(out[1][1], name)), name


@pytest.mark.parametrize("va_max", [0.2, 1.0e30])
@pytest.mark.parametrize("backend", BACKENDS)
def test_boris_uniform_dt(backend, va_max):
    grid, params, f = configs.uniform(8)
    params.boris = 1
    params.va_max = va_max
    sim = make(backend, grid, params, f)
    sim.boundary_conditions()
    sim.eta_calc()
    dt, _ = sim.step(1)
    close(sim)
    rho, e = 1.3, 0.7
    b2 = 0.3 ** 2 + 0.2 ** 2 + 0.1 ** 2
    p = (GAMMA - 1.0) * rho * e
    w2 = b2 / rho
    gb = 1.0 / (1.0 + w2 / va_max ** 2) if w2 >= va_max ** 2 else 1.0
    dx = 1.0 / 8
    want = 0.9 * min(dx / math.sqrt((GAMMA * p + gb * b2) / rho), dx / 0.11)
    print(f"va_max {va_max}: gamma_b {gb:.6f} dt {dt:.15e} expected {want:.15e}")
    assert abs(dt - want) < 1e-13 * want
    if va_max > 1.0:
        grid, params, f = configs.uniform(8)
        sim = make(backend, grid, params, f)
        sim.boundary_conditions()
        sim.eta_calc()
        assert sim.step(1)[0] == dt
        close(sim)


@pytest.mark.parametrize("alpha_av", [1.0, 0.05])
@pytest.mark.parametrize("backend", BACKENDS)
def test_cooling_removes_the_viscous_heating(backend, alpha_av):
    rng = np.random.default_rng(4)
    n = 8
    grid = Grid(n, n, n, ((0.0, 1.0),) * 3, (BC_PERIODIC,) * 6)
    f = Fields(n, n, n)
    f.rho[...] = 1.0 + 0.3 * rng.random(f.rho.shape)
    f.energy[...] = 2.0 + rng.random(f.energy.shape)
    for v in ("vx", "vy", "vz"):
        getattr(f, v)[...] = 0.2 * rng.standard_normal(f.vx.shape)
    res = {}
    for cool in (0, 1):
        params = default_params(visc1=0.2, visc2=0.8, cooling_term=cool, alpha_av=alpha_av)
        sim = make(backend, grid, params, f)
        sim.boundary_conditions()
        sim.set_scalars(dt=1.0e-3)
        sim.lagrangian_step()
        res[cool] = sim.fields()
        close(sim)
    on, off = res[1], res[0]
    for name in ("rho", "vx", "vy", "vz", "visc_dep"):
        assert np.array_equal(getattr(on, name), getattr(off, name)), name
    vd = on.interior("visc_dep")
    assert vd.min() > 0.0
    want = alpha_av * vd / on.interior("rho")
    got = off.interior("energy") - on.interior("energy")
    print(f"alpha_av {alpha_av}: removed {got.mean():.6e}, expected {want.mean():.6e}")
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=0)
