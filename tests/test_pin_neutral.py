"""Pins of the partially ionised plasma (neutral.f90) and its uses in the
step: the neutral fraction, the ionisation energy in the pressure, the
conduction temperature and the Cowling resistivity.

The neutral fraction is a Saha-type equilibrium: with q = r_bar rho bof / 4
(neutral.f90:151-153), xi solves xi / (1 - xi)^2 = q.  The tests solve that
relation by the other root formula, xi = 1 - 2 / (1 + sqrt(1 + 4 q)), with
numpy's exp, and solve the reference's implicit temperature
T = (gamma-1) (e - (1 - xi(T)) ionise_pot) / (2 - xi(T)) with scipy's brentq
to full precision; the bisection under test stops at a 1e-8 step, which
bounds the difference.  The SI form of the same equilibrium, computed from
the physical constants, pins the normalisation (normalise.f90:36-49).
In the step: a pressure profile with the ionisation energy must accelerate
the gas by the gradient of (gamma-1) rho (e - (1 - xi) ionise_pot); a
conduction mode at the ionisation transition must decay at the rate set by
dT/de of the partially ionised gas (not (gamma-1)/(2 - xi)); a current
across B must diffuse at eta + eta_perp and one along B at eta alone.
"""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

from lare3d_amd import BC_PERIODIC, BC_USER
from lare3d_amd.abi import EOS_ION, EOS_PI, F_TBOUNDS
from lare3d_amd.configs import default_params
from lare3d_amd.neutral import (HP_SI, IONISE_POT_SI, KB_SI, ME_SI, MH_SI, MU0_SI, T_RAD_SI, neutral_constants,
                                set_neutral)
from tests.backends import BACKENDS, close, make, start
from tests.physics import along, laplacian_eig, line_problem, rkl2_amplification

MF = 1.2


def _norm(temp_norm, rho_norm=1.0e-6):
    """B_norm giving the normalising temperature temp_norm at rho_norm."""
    return math.sqrt(temp_norm * KB_SI / (MF * MH_SI) * MU0_SI * rho_norm)


def saha_xi(p, T, rho):
    """xi with xi / (1 - xi)^2 = r_bar rho bof / 4, by the root
    1 - 2 / (1 + sqrt(1 + 4 q))."""
    T = np.asarray(T, dtype=float)
    with np.errstate(over="ignore"):  # q = inf (cold): xi = 1
        bof = 1.0 / (0.5 * p.f_bar * p.t_rad * np.sqrt(T)) * np.exp((0.25 * (T / p.t_rad - 1.0) + 1.0) * p.t_bar / T)
        q = p.r_bar * rho * bof / 4.0
        return 1.0 - 2.0 / (1.0 + np.sqrt(1.0 + 4.0 * q))


def eos_temperature(p, e, rho, ip):
    """The root of T = (gamma-1) (e - (1 - xi(T)) ip) / (2 - xi(T))."""
    gm1 = p.gamma - 1.0
    f = lambda T: T - gm1 * (e - (1.0 - saha_xi(p, T, rho)) * ip) / (2.0 - saha_xi(p, T, rho))
    lo = gm1 * max((e - ip) / 2.0, 1e-300)
    return brentq(f, lo, gm1 * e, xtol=1e-15, rtol=1e-15, maxiter=500)


def test_exp_within_an_ulp_of_libm():
    from oracle import oracle
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(-740.0, 709.0, 20000), rng.uniform(-3.0, 3.0, 5000),
                         [0.0, -0.0, 1e-300, 0.5, -0.5, 1.0, 22.7, 709.7, -708.3]])
    worst = 0.0
    for x in xs:
        a, b = oracle.exp(float(x)), math.exp(float(x))
        worst = max(worst, abs(a - b) / math.ulp(b))
    print(f"exp: worst {worst} ulp over {len(xs)} arguments")
    assert worst <= 1.0
    assert oracle.exp(float("inf")) == math.inf and oracle.exp(-math.inf) == 0.0 and oracle.exp(711.0) == math.inf
    assert math.isnan(oracle.exp(float("nan")))


def test_saha_constants_and_si_form():
    """f_bar is the textbook Saha constant (2 pi m_e k_B / h^2)^1.5 =
    2.4147e21 m^-3 K^-1.5, t_bar temp_norm the hydrogen ionisation
    temperature chi / k_B = 157,800 K; the normalised equilibrium equals the
    SI one, xi / (1 - xi)^2 = n / S with n = rho / (mf m_H) and
    S = W (2 pi m_e k_B / h^2)^1.5 T_r sqrt(T) exp(-(0.25 (T/T_r - 1) + 1) chi / (k_B T))."""
    from oracle import oracle
    p = default_params()
    rho_norm, temp_norm = 1.0e-6, 2.0e4
    tn = set_neutral(p, _norm(temp_norm, rho_norm), 10.0, rho_norm, MF, EOS_ION)
    assert abs(tn / temp_norm - 1.0) < 1e-12
    assert abs(p.f_bar / 2.4147e21 - 1.0) < 2e-4
    assert abs(p.t_bar * tn / 157800.0 - 1.0) < 1e-3
    assert abs(p.t_rad * tn - T_RAD_SI) < 1e-9
    # ionise_pot in units of energy0 mbar equals t_bar (chi / (k_B temp_norm))
    assert abs(p.ionise_pot / p.t_bar - 1.0) < 1e-12
    c = neutral_constants(_norm(temp_norm, rho_norm), 10.0, rho_norm, MF, EOS_PI)
    assert c["ionise_pot"] == 0.0
    worst = 0.0
    for T_K in (4000.0, 7000.0, 9000.0, 12000.0, 20000.0):
        for rho_si in (1e-9, 1e-7, 1e-5):
            sf = (2.0 * math.pi * ME_SI * KB_SI / HP_SI ** 2) ** 1.5
            arg = (0.25 * (T_K / T_RAD_SI - 1.0) + 1.0) * IONISE_POT_SI / (KB_SI * T_K)
            S = 0.5 * sf * T_RAD_SI * math.sqrt(T_K) * math.exp(-arg)
            q = rho_si / (MF * MH_SI) / S
            xi_si = 1.0 - 2.0 / (1.0 + math.sqrt(1.0 + 4.0 * q))
            xi = oracle.saha_neutral(p, T_K / tn, rho_si / rho_norm)
            worst = max(worst, abs(xi - xi_si))
            assert abs(xi - xi_si) <= 1e-12 + 1e-10 * xi_si, (T_K, rho_si, xi, xi_si)
    print(f"normalised vs SI equilibrium: worst |dxi| {worst:.2e}")


def test_saha_limits_and_monotonicity():
    """Neutral when cold, ionised when hot; more neutral at higher density
    (fewer free states per particle); and the equilibrium identity
    xi / (1 - xi)^2 = q."""
    from oracle import oracle
    p = default_params()
    set_neutral(p, _norm(2.0e4), 10.0, 1.0e-6, MF, EOS_PI)
    Ts = np.geomspace(0.05, 20.0, 200)
    for rho in (0.01, 1.0, 100.0):
        xi = np.array([oracle.saha_neutral(p, T, rho) for T in Ts])
        assert np.all(np.diff(xi) <= 0.0)
        assert xi[0] > 0.999 and xi[-1] < 0.01
        ref = saha_xi(p, Ts, rho)
        assert np.max(np.abs(xi - ref)) <= 1e-12
    for T in (0.3, 0.5, 0.8):
        xs = [oracle.saha_neutral(p, T, r) for r in (0.01, 0.1, 1.0, 10.0)]
        assert all(b >= a for a, b in zip(xs, xs[1:]))


def _random_state(eos, n=6, seed=3):
    grid, params, f = line_problem(n, 0, ntrans=5, bc_axis=BC_USER, bc_trans=BC_PERIODIC, gamma=5.0 / 3.0)
    set_neutral(params, _norm(4.0e4), 10.0, 1.0e-6, MF, eos)
    rng = np.random.default_rng(seed)
    f.rho[...] = 10.0 ** rng.uniform(-1.0, 1.0, f.rho.shape)
    # temperatures 0.1 .. 0.6 across the transition; energy from the EOS
    ip = params.ionise_pot
    T = rng.uniform(0.1, 0.6, f.energy.shape)
    xi = saha_xi(params, T, f.rho)
    f.energy[...] = T * (2.0 - xi) / (params.gamma - 1.0) + (1.0 - xi) * ip
    return grid, params, f


@pytest.mark.parametrize("eos", [EOS_PI, EOS_ION])
@pytest.mark.parametrize("backend", BACKENDS)
def test_neutral_fraction_solves_the_eos(backend, eos):
    """Every cell of the box (ghosts included): xi_n is the equilibrium at
    the temperature that solves the EOS, to the bisection's 1e-8 step."""
    grid, params, f = _random_state(eos)
    sim = make(backend, grid, params, f)
    try:
        sim.neutral_fraction()
        xi, _ = sim.neutral()
    finally:
        close(sim)
    worst, worst_rel = 0.0, 0.0
    ip = params.ionise_pot
    for idx in np.ndindex(f.energy.shape):
        e, rho = float(f.energy[idx]), float(f.rho[idx])
        T = eos_temperature(params, e, rho, ip)
        x0 = float(saha_xi(params, T, rho))
        slope = abs(float(saha_xi(params, T * (1 + 1e-6), rho)) - x0) / (T * 1e-6)
        d = abs(xi[idx] - x0)
        worst = max(worst, d)
        worst_rel = max(worst_rel, d / (3e-8 * (slope + 1e-3)))
        assert d <= 3e-8 * (slope + 1e-3), (idx, xi[idx], x0, slope)
    print(f"eos {eos}: max |xi - xi(T*)| {worst:.2e} ({worst_rel:.2f} of the bisection bound), "
          f"xi {xi.min():.3f}..{xi.max():.3f}")
    assert xi.min() < 0.2 and xi.max() > 0.8


@pytest.mark.parametrize("backend", BACKENDS)
def test_ionisation_energy_in_the_pressure(backend):
    """EOS_ION, v = 0, no viscosity: one Lagrangian step changes v along a
    by -dt / rho (p(s+1) - p(s)) / dx with p = (gamma-1) rho (e - (1 - xi)
    ionise_pot) of the cells either side of node s (lagran.F90:93-94,
    :244); the ionisation term is a large part of the force."""
    n, a = 32, 1
    grid, params, f = line_problem(n, a, bc_axis=BC_PERIODIC, gamma=5.0 / 3.0, visc1=0.0, visc2=0.0)
    set_neutral(params, _norm(4.0e4), 10.0, 1.0e-6, MF, EOS_ION)
    ip, gm1 = params.ionise_pot, params.gamma - 1.0
    rng = np.random.default_rng(17)
    Tp = 0.3 + 0.05 * np.cos(2.0 * math.pi * np.arange(n) / n + 0.3) + 0.01 * rng.standard_normal(n)
    rho0 = 1.0
    xi_p = saha_xi(params, Tp, rho0)
    ep = Tp * (2.0 - xi_p) / gm1 + (1.0 - xi_p) * ip
    idx = np.arange(-1, n + 3) % n        # cells -1..n+2 -> profile index (cell s <-> s-1 mod n)
    f.energy[...] = along(f.energy.shape, a, ep[(idx - 1) % n])
    f.rho[...] = rho0
    f.bx[...] = 0.0
    sim = make(backend, grid, params, f)
    try:
        start(sim)
        sim.neutral_fraction()
        xi, _ = sim.neutral()
        f0 = sim.fields()
        sim.lagrangian_step()
        dt = sim.scalars().dt
        f1 = sim.fields()
    finally:
        close(sim)
    xi_c = np.moveaxis(xi, a, 0)[2:n + 2, 2, 2]           # cells 1..n
    e_c = np.moveaxis(f0.energy, a, 0)[2:n + 2, 2, 2]
    p = gm1 * rho0 * (e_c - (1.0 - xi_c) * ip)
    p_ideal = gm1 * rho0 * e_c
    dx = 1.0 / n
    # node s (1..n) lies between cells s and s+1 (cell n+1 = cell 1)
    expect = -dt / rho0 * (np.roll(p, -1) - p) / dx
    ideal = -dt / rho0 * (np.roll(p_ideal, -1) - p_ideal) / dx
    name = ("vx", "vy", "vz")[a]
    d = np.moveaxis(f1.interior(name) - f0.interior(name), a, 0)[:, 0, 0]
    err = np.abs(d - expect).max()
    print(f"ionisation pressure: max dv {np.abs(expect).max():.3e}, error {err:.1e}, "
          f"ideal-gas force would differ by {np.abs(ideal - expect).max():.3e}")
    assert err <= 1e-9 * np.abs(expect).max()
    assert np.abs(ideal - expect).max() > 0.2 * np.abs(expect).max()


@pytest.mark.parametrize("backend", BACKENDS)
def test_conduction_at_the_ionisation_transition(backend):
    """EOS_PI, B along x, a small Neumann cosine mode of e about a state at
    the transition: T(e) = (gamma-1) e / (2 - xi(T)) gives the linearised
    conduction rate lam = dT/de kappa T0^2.5 k_d^2 / rho, and one RKL2
    super-step multiplies the mode by R_s(-lam dt).  dT/de here is well
    below the frozen-xi value (gamma-1)/(2 - xi) (ionisation absorbs part
    of the heat)."""
    n, eps, gam = 32, 1e-3, 5.0 / 3.0
    grid, params, f = line_problem(n, 0, bc_axis=BC_USER, bc_trans=BC_USER, gamma=gam, conduction=1,
                                   kappa_0=1e-2, heat_flux_limiter=0)
    set_neutral(params, _norm(1.4e4), 10.0, 1.0e-6, MF, EOS_PI)
    gm1 = gam - 1.0
    f.bx[...] = 1.0
    rho0 = 1.0
    f.rho[...] = rho0
    T0 = 1.0
    xi0 = float(saha_xi(params, T0, rho0))
    e0 = T0 * (2.0 - xi0) / gm1
    k = 6.0 * math.pi
    xc = grid.xc[0]
    f.energy[...] = along(f.energy.shape, 0, e0 * (1.0 + eps * np.cos(k * xc)))
    h = 1e-5 * e0
    dTde = (eos_temperature(params, e0 + h, rho0, 0.0) - eos_temperature(params, e0 - h, rho0, 0.0)) / (2 * h)
    dx = 1.0 / n
    lam_unit = params.kappa_0 * T0 ** 2.5 * laplacian_eig(k, dx) / rho0
    dtp = 0.9 * (dx * dx / (2.0 * params.kappa_0 * 1.05 ** 2.5)) / math.sqrt(3.0)
    dt = 6.0 * dtp
    sim = make(backend, grid, params, f)
    try:
        start(sim)
        sim.neutral_fraction()
        sim.set_scalars(dt=dt)
        sim.conduct_heat()
        s = sim.scalars().n_s_stages
        out = sim.fields()
    finally:
        close(sim)
    en = out.interior("energy")[:, 0, 0]
    x = xc[2:n + 2]
    amp = 2.0 / n * float(np.sum((en - e0) * np.cos(k * x))) / (eps * e0)
    R = rkl2_amplification(s, -dTde * lam_unit * dt)
    R_frozen = rkl2_amplification(s, -gm1 / (2.0 - xi0) * lam_unit * dt)
    print(f"xi0 {xi0:.3f}, dT/de {dTde:.4f} (frozen xi {gm1 / (2.0 - xi0):.4f}), s {s}: mode {amp:.6f}, "
          f"R_s {R:.6f}, frozen-xi R_s {R_frozen:.6f}")
    assert 0.2 < xi0 < 0.8 and s >= 3
    assert abs(amp - R) < 1e-2 * abs(1.0 - R)
    assert abs(R_frozen - R) > 0.2 * abs(1.0 - R)


def _cowling_mode(backend, parallel, n, eps=1e-4, eta=2e-3, b0=0.7):
    """B0 along z with a perturbation along x.  perpendicular: bz = B0 +
    eps cos(kx), so J = -d(bz)/dx y is across B, balanced by the gas
    pressure p0 - B0 dbz - dbz^2 / 2; parallel: by = eps cos(kx), J along
    z = along B (its force is second order).  The perpendicular mode is a
    magnetosonic one: the diffused magnetic pressure leaves a force that
    compresses bz by a fraction ~ (k c dt)^2 / 2 of the diffusion, with dt
    at the CFL limit dx / c, hence the finer grid for it."""
    gam = 5.0 / 3.0
    grid, params, f = line_problem(n, 0, bc_axis=BC_PERIODIC, gamma=gam, resistive_mhd=1, eta_background=eta,
                                   eta0=0.0, j_max=1e30, visc1=0.0, visc2=0.0)
    set_neutral(params, _norm(4.0e4), 10.0, 1.0e-6, MF, EOS_PI, cowling=1)
    k = 2.0 * math.pi
    xc = grid.xc[0]
    d = eps * np.cos(k * xc)
    rho0, p0, gm1 = 1.0, 0.5, gam - 1.0
    f.rho[...] = rho0
    f.bz[...] = b0
    if parallel:
        f.by[...] = along(f.by.shape, 0, d)
        f.energy[...] = p0 / (gm1 * rho0)
    else:
        f.bz[...] = along(f.bz.shape, 0, b0 + d)
        f.energy[...] = along(f.energy.shape, 0, (p0 - b0 * d - 0.5 * d * d) / (gm1 * rho0))
    # eta_perp at the uniform state (neutral.f90:68-136): the reference's
    # vertex field takes bz for the upper bx and by faces (:97-98), so a
    # field B0 along z enters as (B0/2, B0/2, B0), |B|^2 = 1.5 B0^2
    T0 = eos_temperature(params, p0 / (gm1 * rho0), rho0, 0.0)
    xi0 = float(saha_xi(params, T0, rho0))
    bsq = 1.5 * b0 * b0
    # eta_bar set so that eta_perp is 1.5 eta
    params.eta_bar = 1.5 * eta / (xi0 / (1.0 - xi0) * bsq / rho0 ** 2 / math.sqrt(T0))
    ep0 = params.eta_bar * xi0 / (1.0 - xi0) * bsq / rho0 ** 2 / math.sqrt(T0)
    sim = make(backend, grid, params, f)
    try:
        start(sim)
        sim.neutral_fraction()
        sim.perpendicular_resistivity()
        dt, _ = sim.step(1)
        sc = sim.scalars()
        out = sim.fields()
        _, ep = sim.neutral(xi=False, eta_perp=True)
    finally:
        close(sim)
    comp = out.interior("by" if parallel else "bz")[:, 0, 0]
    base = 0.0 if parallel else b0
    amp = 2.0 / n * float(np.sum((comp - base) * np.cos(k * xc[2:n + 2]))) / eps
    lam = laplacian_eig(k, 1.0 / n)
    return amp, sc, dt, ep0, ep, lam, grid


@pytest.mark.parametrize("parallel", [False, True])
@pytest.mark.parametrize("backend", BACKENDS)
def test_cowling_resistivity(backend, parallel):
    """A current across B decays at eta + eta_perp per substep, one along B
    at eta; set_dt takes the substep length from max(eta, eta_perp)
    (lagran.F90:1132-1134)."""
    eta = 2e-3
    amp, sc, dt, ep0, ep, lam, grid = _cowling_mode(backend, parallel, 32 if parallel else 128, eta=eta)
    ns = sc.last_substeps
    dts = dt / ns
    rate = eta if parallel else eta + ep0
    expect = (1.0 - rate * lam * dts) ** ns
    other = (1.0 - (eta + ep0 if parallel else eta) * lam * dts) ** ns
    dxl = 1.0 / sum(1.0 / float(grid.dxb[a][3]) ** 2 for a in range(3))
    print(f"{'parallel' if parallel else 'perpendicular'}: eta_perp {ep0:.4e} (nodes {ep[1:-2, 1:-2, 1:-2].min():.4e}"
          f"..{ep[1:-2, 1:-2, 1:-2].max():.4e}), substeps {ns}, mode {amp:.8f} expected {expect:.8f} "
          f"(other rate {other:.8f}); dtr {sc.dtr:.4e} expected {0.9 * 0.2 * dxl / ep0:.4e}")
    inner = ep[1:-2, 1:-2, 1:-2]
    assert np.abs(inner - ep0).max() <= 1e-3 * ep0
    assert abs(sc.dtr - 0.9 * 0.2 * dxl / ep0) <= 1e-3 * sc.dtr
    assert ns == int(dt / sc.dtr) + 1
    assert abs(amp - expect) < 2e-3 * abs(1.0 - expect)
    assert abs(other - expect) > 0.3 * abs(1.0 - expect)


@pytest.mark.parametrize("backend", BACKENDS)
def test_inverted_bracket_is_a_fault(backend):
    """energy = 0 in one cell inverts neutral_fraction's bracket: the
    reference STOPs (neutral.f90:172-175), here a LARE_F_TBOUNDS status."""
    grid, params, f = _random_state(EOS_PI)
    f.energy[3, 4, 2] = 0.0
    sim = make(backend, grid, params, f)
    try:
        with pytest.raises(Exception) as ei:
            sim.neutral_fraction()
        assert getattr(ei.value, "code", None) == F_TBOUNDS, ei.value
    finally:
        close(sim)
