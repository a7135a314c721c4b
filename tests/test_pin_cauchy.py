"""Pins of the CAUCHY predictor (lagran.F90:285-348 and :945-1048).

With CAUCHY the half-step field is the cell-centred B1 advanced by the
Cauchy solution B1 <- (B1 + (B1 . grad) v dt/2) / (1 + div v dt/2), and the
predictor's J x B is formed from B1 cv1 averaged over the cells around a
node.  A uniform field B0 along axis a with a divergence-free shear, the
velocity component c != a varying along a only, makes this exact: div v = 0,
so cv1 = cv and the pressure stays uniform; B1_c = B0 dt/2 dv_c/dx_a in each
cell (one-sided differences of the node values).  The forces are the
tension B0 dB1_c/dx_a along c and the magnetic pressure along a, so one
Lagrangian step changes v_c by
    dt^2/2 B0^2/rho (v_c(s+1) - 2 v_c(s) + v_c(s-1)) / dx^2
(the discrete Alfven operator), v_a by
    -dt/rho (B1_c(s+1)^2 - B1_c(s)^2) / (2 dx)
(B1_c of the cells s+1 and s either side of node s) and leaves the third
component alone.  The six (a, c) pairs exercise the six transverse
derivatives of the Cauchy update and the six terms of the currents.
"""
import numpy as np
import pytest

from lare3d_amd import BC_PERIODIC
from tests.backends import BACKENDS, close, make
from tests.physics import along, line_problem

PAIRS = [(a, c) for a in range(3) for c in range(3) if a != c]


def _shear_step(backend, a, c, n=32, cauchy=1):
    rho0, p0, b0, gam = 1.3, 0.5, 0.8, 5.0 / 3.0
    grid, params, f = line_problem(n, a, bc_axis=BC_PERIODIC, gamma=gam, visc1=0.0, visc2=0.0)
    params.cauchy = cauchy
    rng = np.random.default_rng(11 + 3 * a + c)
    # node profile along a, periodic: node s and s + n hold the same value
    prof = 0.2 * rng.standard_normal(n)
    idx = np.arange(-2, n + 3) % n
    vname = ("vx", "vy", "vz")[c]
    getattr(f, vname)[...] = along(f.vx.shape, a, prof[idx])
    f.rho[...] = rho0
    f.energy[...] = p0 / ((gam - 1.0) * rho0)
    getattr(f, ("bx", "by", "bz")[a])[...] = b0
    sim = make(backend, grid, params, f)
    try:
        sim.boundary_conditions()
        f0 = sim.fields()
        sim.lagrangian_step()
        dt = sim.scalars().dt
        f1 = sim.fields()
    finally:
        close(sim)
    dx = 1.0 / n
    lap = np.roll(prof, -1) - 2.0 * prof + np.roll(prof, 1)
    expect = 0.5 * dt * dt * b0 * b0 / rho0 * lap / dx ** 2
    # B1_c of cell s (between nodes s-1 and s), and the pressure term at node s
    b1 = b0 * 0.5 * dt * (prof - np.roll(prof, 1)) / dx
    press = -dt / rho0 * (np.roll(b1, -1) ** 2 - b1 ** 2) / (2.0 * dx)
    dv = {}
    for q, nm in enumerate(("vx", "vy", "vz")):
        d = np.moveaxis(f1.interior(nm) - f0.interior(nm), a, 0)
        dv[q] = d
    # interior nodes 1..n along a are the profile's nodes 1..n (index s mod n)
    nodes = np.arange(1, n + 1) % n
    return expect[nodes], press[nodes], dv, dt


@pytest.mark.parametrize("pair", PAIRS)
@pytest.mark.parametrize("backend", BACKENDS)
def test_cauchy_shear_is_the_discrete_alfven_operator(backend, pair):
    a, c = pair
    expect, press, dv, dt = _shear_step(backend, a, c)
    err = np.abs(dv[c][:, 0, 0] - expect).max()
    errp = np.abs(dv[a][:, 0, 0] - press).max()
    print(f"B along {a}, v_{c}: dt {dt:.4e}, max change {np.abs(expect).max():.3e} (error {err:.1e}), "
          f"along B {np.abs(press).max():.3e} (error {errp:.1e})")
    assert np.abs(expect).max() > 1e-6 and np.abs(press).max() > 1e-6
    assert err <= 1e-9 * np.abs(expect).max()
    assert errp <= 1e-9 * np.abs(press).max()
    # uniform across the transverse axes; the third component unchanged
    for q in (a, c):
        assert np.abs(dv[q] - dv[q][:, :1, :1]).max() <= 1e-15
    assert np.abs(dv[3 - a - c]).max() <= 1e-15


def _compression_step(backend, a, c, n=32):
    """Compression along a (v_a varying along a) of a uniform field B0 along
    c: the Cauchy update gives B1_c = B0 / (1 + dv) with dv = dt/2 dv_a/dx_a,
    so B1_c cv1 = B0 cv; the predictor pressure is p1 = e1 (gamma - 1) rho
    cv / cv1 with e1 = e - p (cv1/cv - 1) / rho (lagran.F90:234-249)."""
    rho0, p0, b0, gam = 1.3, 0.5, 0.8, 5.0 / 3.0
    grid, params, f = line_problem(n, a, bc_axis=BC_PERIODIC, gamma=gam, visc1=0.0, visc2=0.0)
    params.cauchy = 1
    rng = np.random.default_rng(23 + 3 * a + c)
    prof = 0.2 * rng.standard_normal(n)
    idx = np.arange(-2, n + 3) % n
    getattr(f, ("vx", "vy", "vz")[a])[...] = along(f.vx.shape, a, prof[idx])
    f.rho[...] = rho0
    e0 = p0 / ((gam - 1.0) * rho0)
    f.energy[...] = e0
    getattr(f, ("bx", "by", "bz")[c])[...] = b0
    sim = make(backend, grid, params, f)
    try:
        sim.boundary_conditions()
        f0 = sim.fields()
        sim.lagrangian_step()
        dt = sim.scalars().dt
        f1 = sim.fields()
    finally:
        close(sim)
    dx = 1.0 / n
    dv = (prof - np.roll(prof, 1)) / dx * (0.5 * dt)       # cell s, between nodes s-1 and s
    ratio = 1.0 + dv                                       # cv1 / cv
    p1 = (e0 - p0 * (ratio - 1.0) / rho0) * (gam - 1.0) * rho0 / ratio
    bc = b0 / ratio                                        # the compressed field, B0 cv / cv1
    f_gas = -(np.roll(p1, -1) - p1) / dx                   # node s between cells s and s+1
    f_mag = -(np.roll(bc, -1) - bc) / dx * 2.0 * b0 / (ratio + np.roll(ratio, -1))
    expect = dt * (f_gas + f_mag) / rho0
    d = np.moveaxis(f1.interior(("vx", "vy", "vz")[a]) - f0.interior(("vx", "vy", "vz")[a]), a, 0)
    nodes = np.arange(1, n + 1) % n
    return d, expect[nodes], (dt * f_mag / rho0)[nodes]


@pytest.mark.parametrize("pair", PAIRS)
@pytest.mark.parametrize("backend", BACKENDS)
def test_cauchy_compression_of_a_transverse_field(backend, pair):
    a, c = pair
    d, expect, mag = _compression_step(backend, a, c)
    err = np.abs(d[:, 0, 0] - expect).max()
    print(f"compression along {a}, B along {c}: max change {np.abs(expect).max():.3e}, of it magnetic "
          f"{np.abs(mag).max():.3e}; max error {err:.1e}")
    assert np.abs(mag).max() > 0.1 * np.abs(expect).max()
    assert err <= 1e-9 * np.abs(expect).max()
    assert np.abs(d - d[:, :1, :1]).max() <= 1e-15
