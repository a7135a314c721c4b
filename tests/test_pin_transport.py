"""Pins of conduction, resistivity, energy diagnostics, boundary fills and
the reference's documented quirks against closed forms.

RKL2: one super-step on a small-amplitude Neumann cosine mode must multiply
the mode by the scheme's closed-form stability polynomial
R_s(z) = a_s + b_s P_s(1 + w1 z), P_s Legendre (Meyer, Balsara & Aslam 2012), with z the
discrete parallel-diffusion rate times dt; an oblique field must scale the
rate by cos^2 of the angle (anisotropic projection of heat_flux).  The
stage count must follow the closed form of the explicit parabolic step.
Resistivity: a By mode must decay by (1 - eta lambda dt_sub)^substeps and
heat by eta sum J^2 dV dt.  energy_account: closed forms for uniform fields.
"""
import math

import numpy as np
import pytest

from lare3d_amd import BC_PERIODIC, BC_USER, configs
from tests.backends import BACKENDS, close, make, run_until, start
from tests.physics import along, laplacian_eig, line_problem, rkl2_amplification


def _cond_mode(backend, theta_deg=0.0, n=32, eps=1e-7, s_target=7, m=1):
    gam = 5.0 / 3.0
    grid, params, f = line_problem(n, 0, bc_axis=BC_USER, bc_trans=BC_USER, gamma=gam, conduction=1,
                                   kappa_0=1e-2, heat_flux_limiter=0)
    th = math.radians(theta_deg)
    f.bx[...] = math.cos(th)
    f.by[...] = math.sin(th)
    f.rho[...] = 1.0
    tfac = (gam - 1.0) / 2.0
    L = 1.0
    k = math.pi * m / L
    xc = grid.xc[0]
    T = 1.0 + eps * np.cos(k * xc)
    f.energy[...] = along(f.energy.shape, 0, T / tfac)
    dx = 1.0 / n
    # explicit parabolic step (larsen = 1) and the dt giving s_target stages
    dtp = 0.9 * (dx * dx / (2.0 * params.kappa_0 * (1.0 + eps) ** 2.5)) / math.sqrt(3.0)
    r = ((2 * s_target + 1 - 0.5) ** 2 - 9.0) / 16.0
    dt = r * dtp
    sim = make(backend, grid, params, f)
    start(sim)
    sim.set_scalars(dt=dt)
    sim.conduct_heat()
    s = sim.scalars().n_s_stages
    out = sim.fields()
    close(sim)
    Tn = out.interior("energy")[:, 0, 0] * tfac
    x = xc[2:n + 2]
    amp = 2.0 / n * float(np.sum((Tn - 1.0) * np.cos(k * x))) / eps
    lam = tfac * params.kappa_0 * 1.0 * laplacian_eig(k, dx) * math.cos(th) ** 2
    expect_s = math.ceil(0.5 * (math.sqrt(9.0 + 16.0 * dt / dtp) - 1.0))
    expect_s += 1 - expect_s % 2
    return amp, rkl2_amplification(s, -lam * dt), s, expect_s, lam * dt


@pytest.mark.parametrize("backend", BACKENDS)
def test_rkl2_legendre_polynomial(backend):
    amp, R, s, es, z = _cond_mode(backend)
    print(f"RKL2 s={s} z={z:.3f}: mode factor {amp:.9f}, R_s(-z) = {R:.9f}")
    assert s == es
    assert abs(amp - R) < 1e-5 * abs(1.0 - R)


@pytest.mark.parametrize("s_target", [3, 11, 21])
def test_rkl2_stage_counts(s_target):
    amp, R, s, es, z = _cond_mode("oracle", s_target=s_target, m=3)
    assert s == es
    assert abs(amp - R) < 1e-5 * max(abs(1.0 - R), 1e-3)


def test_conduction_oblique_field():
    amp, R, s, es, z = _cond_mode("oracle", theta_deg=30.0)
    assert abs(amp - R) < 1e-5 * abs(1.0 - R)
    amp0, R0, _, _, _ = _cond_mode("oracle", theta_deg=0.0)
    assert abs(amp0 - amp) > 10 * abs(amp - R)  # the projection matters


def _resistive_mode(backend, n=32, eps=1e-3, eta=2e-3):
    gam = 5.0 / 3.0
    grid, params, f = line_problem(n, 0, bc_axis=BC_PERIODIC, gamma=gam, resistive_mhd=1,
                                   eta_background=eta, eta0=0.0, j_max=1e30, visc1=0.0, visc2=0.0)
    k = 2.0 * math.pi
    xc = grid.xc[0]
    f.by[...] = along(f.by.shape, 0, eps * np.cos(k * xc))
    f.rho[...] = 1.0
    f.energy[...] = 1.0 / ((gam - 1.0))
    sim = make(backend, grid, params, f)
    start(sim)
    dt, _ = sim.step(1)
    sc = sim.scalars()
    out = sim.fields()
    close(sim)
    by = out.interior("by")[:, 0, 0]
    amp = 2.0 / n * float(np.sum(by * np.cos(k * xc[2:n + 2]))) / eps
    dx = 1.0 / n
    lam = laplacian_eig(k, dx)
    ns = sc.last_substeps
    dts = dt / ns
    expect = (1.0 - eta * lam * dts) ** ns
    kd2 = lam  # (2 sin(k dx/2)/dx)^2
    V = 1.0 * (4 * dx) * (4 * dx)
    heat = sum(eta * kd2 * V / 2.0 * (eps * (1.0 - eta * lam * dts) ** q) ** 2 * dts for q in range(ns))
    return amp, expect, sc, dt, heat


@pytest.mark.parametrize("backend", BACKENDS)
def test_resistive_diffusion(backend):
    amp, expect, sc, dt, heat = _resistive_mode(backend)
    print(f"resistive: substeps {sc.last_substeps} mode {amp:.10f} expected {expect:.10f}; "
          f"ohmic {sc.total_ohmic_heating:.6e} expected {heat:.6e}")
    assert abs(amp - expect) < 1e-4 * (1.0 - expect)
    dxl = 1.0 / (1.0 / (1 / 32) ** 2 + 2.0 / (4 / 32 / 4) ** 2)
    assert sc.last_substeps == int(dt / sc.dtr) + 1
    assert abs(sc.total_ohmic_heating - heat) < 1e-3 * heat


@pytest.mark.parametrize("backend", BACKENDS)
def test_resistive_diffusion_rk4(backend):
    """FOURTHORDER (lagran.F90:1329-1415): a By mode of the discrete
    Laplacian's eigenvalue lam is multiplied per substep by the classical
    Runge-Kutta polynomial R4(z) = 1 + z + z^2/2 + z^3/6 + z^4/24,
    z = -eta lam dt_sub; its four stages see the amplitudes 1, 1 + z/2,
    1 + z/2 + z^2/4 and 1 + z + z^2/2 + z^3/4, and the heating is
    eta lam V / 2 a^2 dt_sub / 6 (s1^2 + 2 s2^2 + 2 s3^2 + s4^2).  The
    mode has 4 cells per wavelength, so that |z| ~ 0.1 and a first-order
    step (1 + z) misses by z^2/2 ~ 5e-3."""
    n, eps, eta, gam = 32, 1e-3, 2e-3, 5.0 / 3.0
    grid, params, f = line_problem(n, 0, bc_axis=BC_PERIODIC, gamma=gam, resistive_mhd=1,
                                   eta_background=eta, eta0=0.0, j_max=1e30, visc1=0.0, visc2=0.0)
    params.fourth_order = 1
    k = 2.0 * math.pi * 8
    xc = grid.xc[0]
    f.by[...] = along(f.by.shape, 0, eps * np.cos(k * xc))
    f.rho[...] = 1.0
    f.energy[...] = 1.0 / (gam - 1.0)
    sim = make(backend, grid, params, f)
    start(sim)
    f0 = sim.fields()
    dt, _ = sim.step(1)
    sc = sim.scalars()
    out = sim.fields()
    close(sim)
    by = out.interior("by")[:, 0, 0]
    amp = 2.0 / n * float(np.sum(by * np.cos(k * xc[2:n + 2]))) / eps
    dx = 1.0 / n
    # internal energy gained: the heat (the work of the eps^2 forces is
    # second order in it)
    cv = dx ** 3
    gain = float(np.sum(out.interior("rho") * out.interior("energy") - f0.interior("rho") * f0.interior("energy"))) * cv
    lam = laplacian_eig(k, dx)
    ns = sc.last_substeps
    z = -eta * lam * dt / ns
    r4 = 1.0 + z + z * z / 2.0 + z ** 3 / 6.0 + z ** 4 / 24.0
    expect = r4 ** ns
    stages = (1.0, 1.0 + z / 2.0, 1.0 + z / 2.0 + z * z / 4.0, 1.0 + z + z * z / 2.0 + z ** 3 / 4.0)
    weight = (stages[0] ** 2 + 2.0 * stages[1] ** 2 + 2.0 * stages[2] ** 2 + stages[3] ** 2) / 6.0
    V = 1.0 * (4 * dx) * (4 * dx)
    heat = sum(eta * lam * V / 2.0 * (eps * r4 ** q) ** 2 * weight * dt / ns for q in range(ns))
    print(f"rk4 resistive: substeps {ns}, z {z:.4f}: mode {amp:.10f} expected {expect:.10f} "
          f"(first order {(1.0 + z) ** ns:.10f}); ohmic {sc.total_ohmic_heating:.6e} expected {heat:.6e}, "
          f"internal energy gained {gain:.6e}")
    assert abs(z) > 0.05
    assert abs(amp - expect) < 1e-3 * abs(expect - (1.0 + z) ** ns)
    assert abs(sc.total_ohmic_heating - heat) < 1e-3 * heat
    assert abs(gain - heat) < 1e-4 * heat


def test_eta_threshold():
    """eta = eta_bg + eta0 where |J| > j_max (lagran.F90:1218)."""
    n = 8
    grid, params, f = configs.uniform(n, resistive_mhd=1, eta_background=1e-3, eta0=5e-2, j_max=0.5)
    f.bx[...] = 0.0; f.by[...] = 0.0; f.bz[...] = 0.0
    # by = c * x_c  ->  jz = dby/dx = c everywhere the stencil sees the ramp
    for c, expect in ((0.4, 1e-3), (0.6, 1e-3 + 5e-2)):
        xc = grid.xc[0]
        f.by[...] = along(f.by.shape, 0, c * xc)
        sim = make("oracle", grid, params, f)
        sim.eta_calc()
        eta = sim.fields().eta[3:6, 3:6, 3:6]
        assert np.allclose(eta, expect, rtol=0, atol=1e-15)
        close(sim)


@pytest.mark.parametrize("backend", BACKENDS)
def test_energy_account_uniform(backend):
    grid, params, f = configs.uniform(8)
    sim = make(backend, grid, params, f)
    start(sim)
    eb, eke, eint = sim.energy_account()
    close(sim)
    V = 1.0
    assert abs(eb - 0.5 * (0.3 ** 2 + 0.2 ** 2 + 0.1 ** 2) * V) < 1e-13
    assert abs(eke - 0.5 * 1.3 * (0.11 ** 2 + 0.07 ** 2 + 0.05 ** 2) * V) < 1e-13
    assert abs(eint - 1.3 * 0.7 * V) < 1e-13


def test_step_one_has_no_shock_heating():
    """dt is 0 on the first step, so every edge viscosity vanishes (Q1)."""
    grid, params, f = configs.harris(12)
    sim = make("oracle", grid, params, f)
    start(sim)
    sim.step(1)
    v1 = sim.fields().visc_dep.copy()
    sim.step(2)
    v2 = sim.fields().visc_dep
    assert np.all(v1 == 0.0)
    assert float(np.abs(v2).max()) > 0.0 and float(v2.min()) >= 0.0


def test_stage_cap_rescales_dt():
    """Stage counts >= 60 rescale dt by (2*60^2 - 9)/(2 s^2 - 9) (lagran.F90:1153-1162)."""
    grid, params, f = configs.uniform(8, conduction=1, kappa_0=5000.0, heat_flux_limiter=0)
    f.vx[...] = 0.0; f.vy[...] = 0.0; f.vz[...] = 0.0
    sim = make("oracle", grid, params, f)
    start(sim)
    dt, _ = sim.step(1)
    gam, rho, e = params.gamma, 1.3, 0.7
    p = (gam - 1.0) * rho * e
    dx = 1.0 / 8
    cfl = 0.9 * dx / math.sqrt((gam * p + 0.14) / rho)
    T = (gam - 1.0) / 2.0 * e
    dtp = 0.9 * (rho * dx * dx / (2.0 * 5000.0 * T ** 2.5)) / math.sqrt(3.0)
    s = math.ceil(0.5 * (math.sqrt(9.0 + 16.0 * cfl / dtp) - 1.0))
    s += 1 - s % 2
    assert s >= 60
    assert abs(dt - cfl * (2 * 3600 - 9) / (2 * s * s - 9)) < 1e-14 * dt
    close(sim)


# ---------------------------------------------------------------------------
# boundary fills versus the slab rules of mpiboundary.f90 / boundary.f90,
# implemented independently with numpy slicing
# ---------------------------------------------------------------------------
def _slab_reference(a, lo, n, per, kind, comp=None):
    """Apply x, y, z exchanges (periodic axes) then x, y, z user fills."""
    a = a.copy()
    def sl(ax, idx):
        s = [slice(None)] * 3
        s[ax] = idx - lo[ax]
        return tuple(s)
    for ax in range(3):
        if not per[ax]:
            continue
        N = n[ax]
        if kind == "node":
            for l in range(3): a[sl(ax, N + l)] = a[sl(ax, l)]
            for l in range(2): a[sl(ax, -2 + l)] = a[sl(ax, N - 2 + l)]
        elif kind == "face" and comp == ax:
            for l in range(2): a[sl(ax, N + 1 + l)] = a[sl(ax, 1 + l)]
            for l in range(3): a[sl(ax, -2 + l)] = a[sl(ax, N - 2 + l)]
        else:
            for l in range(2): a[sl(ax, N + 1 + l)] = a[sl(ax, 1 + l)]
            for l in range(2): a[sl(ax, -1 + l)] = a[sl(ax, N - 1 + l)]
    for ax in range(3):
        if per[ax]:
            continue
        N = n[ax]
        if kind == "node":
            for l in (-2, -1, 0, N, N + 1, N + 2): a[sl(ax, l)] = 0.0
        elif kind == "face" and comp == ax:
            a[sl(ax, -1)] = a[sl(ax, 1)]; a[sl(ax, -2)] = a[sl(ax, 2)]
            a[sl(ax, N + 1)] = a[sl(ax, N - 1)]; a[sl(ax, N + 2)] = a[sl(ax, N - 2)]
        else:
            a[sl(ax, 0)] = a[sl(ax, 1)]; a[sl(ax, -1)] = a[sl(ax, 2)]
            a[sl(ax, N + 1)] = a[sl(ax, N)]; a[sl(ax, N + 2)] = a[sl(ax, N - 1)]
    return a


@pytest.mark.parametrize("bcs", [(BC_PERIODIC,) * 6, (BC_USER,) * 6,
                                 (BC_USER, BC_USER, BC_PERIODIC, BC_PERIODIC, BC_USER, BC_USER)])
def test_boundary_slab_rules(bcs):
    n = 6
    grid, params, f = configs.uniform(n, bcs=bcs)
    rng = np.random.default_rng(11)
    for name in ("rho", "energy", "vx", "vy", "vz", "bx", "by", "bz"):
        getattr(f, name)[...] = rng.standard_normal(getattr(f, name).shape)
    sim = make("oracle", grid, params, f)
    sim.boundary_conditions()
    out = sim.fields()
    per = [bcs[2 * a] == BC_PERIODIC for a in range(3)]
    nn = (n, n, n)
    specs = {"rho": ((-1, -1, -1), "cell", None), "energy": ((-1, -1, -1), "cell", None),
             "vx": ((-2, -2, -2), "node", None), "vy": ((-2, -2, -2), "node", None), "vz": ((-2, -2, -2), "node", None),
             "bx": ((-2, -1, -1), "face", 0), "by": ((-1, -2, -1), "face", 1), "bz": ((-1, -1, -2), "face", 2)}
    for name, (lo, kind, comp) in specs.items():
        ref = _slab_reference(getattr(f, name), lo, nn, per, kind, comp)
        assert np.array_equal(getattr(out, name), ref), name
    close(sim)


def test_temperature_powers_correctly_rounded():
    """T**2.5 and T**1.5 of the conduction flux and limiter (conduct.f90:27,
    :83) are evaluated correctly rounded by the oracle (and by the device,
    with the same operations): compare with the exact power rounded once,
    from 50-digit decimal arithmetic, over nine decades of temperature and
    at exact cases (perfect squares, 0)."""
    from decimal import Decimal, localcontext
    from oracle.oracle import lib
    rng = np.random.default_rng(7)
    ts = np.exp(rng.uniform(math.log(1e-6), math.log(1e3), 20000))
    ts = np.concatenate([ts, [0.25, 1.0, 4.0, 9.0, 2.0 ** -20, 1e-300, 1e100, 0.7108961326867776]])
    with localcontext() as ctx:
        ctx.prec = 50
        for t in ts:
            d = Decimal(float(t))
            r = d.sqrt()
            assert lib().orc_pow25(float(t)) == float(d * d * r), t
            assert lib().orc_pow15(float(t)) == float(d * r), t
    assert lib().orc_pow25(0.0) == 0.0 and lib().orc_pow15(0.0) == 0.0
    assert math.isnan(lib().orc_pow25(-1.0)) and math.isinf(lib().orc_pow25(1e200))
