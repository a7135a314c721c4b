"""Synthetic initial conditions of the benchmark configurations.

In the reference these are user files (control.f90 + initial_conditions.f90,
set_initial_conditions at initial_conditions.f90:45-109).  Each function here
returns (grid, params, fields) for one configuration; every physics switch
is set explicitly (the reference's control.f90 defaults switch on viscous
damping, cooling and conduction).  B is set on faces so that its discrete
divergence vanishes (from an edge vector potential where it is not trivially
div-free).
"""
import math

import numpy as np

from .abi import BC_PERIODIC, BC_USER, LareParams, PREC_FP64
from .fields import Fields
from .grid import Grid


def default_params(**kw):
    p = LareParams()
    p.gamma = 5.0 / 3.0
    p.visc1 = 0.1
    p.visc2 = 0.5
    p.eta_background = 0.0
    p.eta0 = 0.0
    p.j_max = 1.0e10
    p.kappa_0 = 0.0
    p.flux_limiter = 0.1
    p.xi_n = 0.0
    p.dt_from_restart = 0.0
    p.resistive_mhd = 0
    p.conduction = 0
    p.heat_flux_limiter = 1
    p.use_viscous_damping = 0
    p.restart = 0
    p.precision = PREC_FP64
    p.alpha_av = 0.05
    p.va_max = 10.0
    p.cooling_term = 0
    p.boris = 0
    p.shock_limiter = 0
    p.shock_compression = 0
    p.rke = 0
    p.cauchy = 0
    p.fourth_order = 0
    p.eos_number = 1
    p.cowling_resistivity = 0
    for k in ("ionise_pot", "f_bar", "t_bar", "r_bar", "eta_bar", "t_rad"):
        setattr(p, k, 0.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _mesh(grid, kinds):
    """Coordinate arrays for a staggering: kinds[a] is 'b' (xb, -2..n+2 for
    nodes / -2 or -1 lower for faces) or 'c' (xc, -1..n+2)."""
    out = []
    for a, k in enumerate(kinds):
        out.append(grid.xb[a] if k == "b" else grid.xc[a])
    return out


def _fill(arr, lo, fn, xs):
    """arr[i,j,k] = fn(x, y, z) over the whole array; xs[a] is indexed by
    (Fortran index - base[a])."""
    X, Y, Z = np.meshgrid(xs[0], xs[1], xs[2], indexing="ij")
    arr[...] = fn(X, Y, Z)


def _coords(grid, kind_per_axis, lower):
    """Coordinates for every array index, given per-axis 'b'/'c' and the
    array's lower bounds."""
    xs = []
    for a in range(3):
        n = grid.n[a]
        idx = np.arange(lower[a], n + 3)
        if kind_per_axis[a] == "b":
            xs.append(np.array([grid.xb[a][i + 2] for i in idx]))
        else:
            # cell centres exist for -1..n+2; index -2 only occurs for the
            # extra low layer of a face array's normal axis, which is 'b'
            xs.append(np.array([grid.xc[a][i + 1] for i in idx]))
    return xs


def orszag_tang(n=64):
    """3D Orszag-Tang vortex, periodic [0, 2pi]^3, ideal MHD (config 1)."""
    tp = 2.0 * math.pi
    grid = Grid(n, n, n, ((0.0, tp),) * 3, (BC_PERIODIC,) * 6)
    params = default_params(gamma=5.0 / 3.0, visc1=0.1, visc2=0.5)
    f = Fields(n, n, n)
    gam = params.gamma
    rho0, p0 = 25.0 / 9.0, 5.0 / 3.0
    f.rho[...] = rho0
    f.energy[...] = p0 / ((gam - 1.0) * rho0)
    xs = _coords(grid, "bbb", (-2, -2, -2))
    _fill(f.vx, None, lambda x, y, z: -np.sin(y), xs)
    _fill(f.vy, None, lambda x, y, z: np.sin(x), xs)
    _fill(f.vz, None, lambda x, y, z: 0.1 * np.sin(z), xs)
    _fill(f.bx, None, lambda x, y, z: -np.sin(y), _coords(grid, "bcc", (-2, -1, -1)))
    _fill(f.by, None, lambda x, y, z: np.sin(2.0 * x), _coords(grid, "cbc", (-1, -2, -1)))
    f.bz[...] = 0.0
    return grid, params, f


def harris(n=64, seed=12345, ny=None, nz=None, noise=0.01, stretch=(None, None, None), damping=None):
    """Resistive Harris sheet (config 2): Bx = tanh(y/0.5),
    rho = 1 + sech^2(y/0.5), p = 0.5 rho (total pressure balance,
    isothermal), eta = 1e-3 plus 1e-2 above |J| = 1.0, seeded velocity noise,
    shock viscosity on.  Periodic in x and z, BC_USER walls in y.
    stretch: stretched length per axis (setup.F90:256-285), or None.
    damping: if set, viscous damping (lagran.F90:483-524) is switched on with
    visc3 = damping * (1 + 0.5 tanh(y) + 0.25 sin(2 pi x / Lx) cos(2 pi z / Lz))
    on the nodes."""
    ny = ny or n
    nz = nz or n
    L = 4.0
    grid = Grid(n, ny, nz, ((-2 * L, 2 * L), (-L, L), (-2 * L, 2 * L)),
                (BC_PERIODIC, BC_PERIODIC, BC_USER, BC_USER, BC_PERIODIC, BC_PERIODIC), stretch)
    params = default_params(resistive_mhd=1, eta_background=1.0e-3, eta0=1.0e-2, j_max=1.0)
    f = Fields(n, ny, nz)
    gam = params.gamma
    w = 0.5
    sech2 = lambda y: 1.0 / np.cosh(y / w) ** 2
    xs = _coords(grid, "ccc", (-1, -1, -1))
    _fill(f.rho, None, lambda x, y, z: 1.0 + sech2(y), xs)
    _fill(f.energy, None, lambda x, y, z: 0.5 * (1.0 + sech2(y)) / ((gam - 1.0) * (1.0 + sech2(y))), xs)
    _fill(f.bx, None, lambda x, y, z: np.tanh(y / w), _coords(grid, "bcc", (-2, -1, -1)))
    rng = np.random.default_rng(seed)
    for name in ("vx", "vy", "vz"):
        a = getattr(f, name)
        a[...] = noise * (2.0 * rng.random(a.shape) - 1.0)
    if damping is not None:
        params.use_viscous_damping = 1
        lx, lz = 4.0 * L, 4.0 * L
        _fill(f.visc3, None, lambda x, y, z: damping * (1.0 + 0.5 * np.tanh(y) + 0.25 * np.sin(2.0 * np.pi * x / lx)
                                                        * np.cos(2.0 * np.pi * z / lz)),
              _coords(grid, "bbb", (-2, -2, -2)))
    return grid, params, f


def arcade_conduction(nx=64, ny=64, nz=128, s_target=21, b0=0.1, shear=0.5):
    """Braginskii conduction on a sheared arcade (config 3): B from an edge
    vector potential A = (0, b0/k cos(kx) e^{-lz}, shear*b0/k cos(kx) e^{-lz}),
    T = 1 + 10 exp(-r^2/0.05^2) about (0, 0, 0.3), rho = 1, kappa_par =
    kappa_0 T^{5/2}; |B|^2 ~ 1e-2 so that min_b = 1e-10 gives
    kappa_perp/kappa_par ~ 1e-8.  x, y periodic; z BC_USER.
    Returns also the fixed dt giving about s_target RKL2 stages."""
    Lx = 1.0
    grid = Grid(nx, ny, nz, ((-0.5, 0.5), (-0.5, 0.5), (0.0, Lx * nz / nx)),
                (BC_PERIODIC, BC_PERIODIC, BC_PERIODIC, BC_PERIODIC, BC_USER, BC_USER))
    params = default_params(conduction=1, kappa_0=1.0e-3, flux_limiter=0.1, heat_flux_limiter=1)
    f = Fields(nx, ny, nz)
    gam = params.gamma
    k = 2.0 * math.pi / Lx
    l = 0.5 * k
    # edge potentials: A_y at (xb_i, yc_j, zb_k); A_z at (xb_i, yb_j, zc_k)
    xb, yb, zb = grid.xb
    xc, yc, zc = grid.xc
    Ay = lambda i, k_: b0 / k * np.cos(k * xb[i + 2]) * np.exp(-l * zb[k_ + 2])
    Az = lambda i, k_: shear * b0 / k * np.cos(k * xb[i + 2]) * np.exp(-l * zc[k_ + 1])
    I = np.arange(-2, nx + 3)
    K = np.arange(-2, nz + 3)
    dxb, dyb, dzb = grid.dxb
    # bx(i,j,k) = (Az(i,j,k) - Az(i,j-1,k))/dyb(j) - (Ay(i,j,k) - Ay(i,j,k-1))/dzb(k)
    # Az independent of j -> first term zero
    for kk in range(-1, nz + 3):
        for ii in range(-2, nx + 3):
            val = -(Ay(ii, kk) - Ay(ii, kk - 1)) / dzb[kk + 1]
            f.bx[ii + 2, :, kk + 1] = val
    # by(i,j,k) = -(Az(i,j,k) - Az(i-1,j,k))/dxb(i)
    for kk in range(-1, nz + 3):
        for ii in range(-1, nx + 3):
            f.by[ii + 1, :, kk + 1] = -(Az(ii, kk) - Az(ii - 1, kk)) / dxb[ii + 1]
    # bz(i,j,k) = (Ay(i,j,k) - Ay(i-1,j,k))/dxb(i)
    for kk in range(-2, nz + 3):
        for ii in range(-1, nx + 3):
            f.bz[ii + 1, :, kk + 2] = (Ay(ii, kk) - Ay(ii - 1, kk)) / dxb[ii + 1]
    f.rho[...] = 1.0
    xs = _coords(grid, "ccc", (-1, -1, -1))
    zc0 = 0.3
    T = lambda x, y, z: 1.0 + 10.0 * np.exp(-(x * x + y * y + (z - zc0) ** 2) / 0.05 ** 2)
    tfac = (gam - 1.0) / (2.0 - params.xi_n)
    _fill(f.energy, None, lambda x, y, z: T(x, y, z) / tfac, xs)
    # dt for about s_target stages: s = (sqrt(9 + 16 dt/dtp) - 1)/2
    tmax = 11.0
    dmin = min(min(grid.dxb[a]) for a in range(3))
    dtp = 0.9 / math.sqrt(3.0) * 1.0 * dmin ** 2 / (2.0 * params.kappa_0 * tmax ** 2.5)
    r = ((2 * s_target - 1 + 1.0) ** 2 - 9.0) / 16.0
    dt = 0.98 * r * dtp
    return grid, params, f, dt


def harris_conduction(n=32, seed=0, s_target=9, kappa_0=1.0e-2, blobs=12):
    """Training and evaluation states of the conduction surrogate: the
    resistive Harris sheet (seeded velocity noise) with Braginskii conduction
    on and a seeded temperature structure filling the box,
    T = T_h (1 + sum_m A_m exp(-|x - x_m|^2 / w_m^2)) over `blobs` Gaussians
    with A_m in [0.5, 3], w_m in [0.5, 2] and x_m uniform in the box.
    Returns also the dt giving about s_target RKL2 stages."""
    grid, params, f = harris(n, seed=seed)
    params.conduction = 1
    params.kappa_0 = kappa_0
    params.flux_limiter = 0.1
    params.heat_flux_limiter = 1
    rng = np.random.default_rng(1000 + seed)
    gam = params.gamma
    tfac = (gam - 1.0) / (2.0 - params.xi_n)
    xs = _coords(grid, "ccc", (-1, -1, -1))
    X, Y, Z = np.meshgrid(xs[0], xs[1], xs[2], indexing="ij")
    bump = np.ones_like(X)
    for _ in range(blobs):
        amp = rng.uniform(0.5, 3.0)
        w = rng.uniform(0.5, 2.0)
        x0 = [rng.uniform(grid.extent[a][0], grid.extent[a][1]) for a in range(3)]
        bump += amp * np.exp(-((X - x0[0]) ** 2 + (Y - x0[1]) ** 2 + (Z - x0[2]) ** 2) / w ** 2)
    f.energy[...] = f.energy * bump
    tmax = float((tfac * f.energy).max())
    dmin = min(min(grid.dxb[a]) for a in range(3))
    rmin = float(f.rho.min())
    dtp = 0.9 / math.sqrt(3.0) * rmin * dmin ** 2 / (2.0 * kappa_0 * tmax ** 2.5)
    r = ((2 * s_target - 1 + 1.0) ** 2 - 9.0) / 16.0
    dt = 0.98 * r * dtp
    return grid, params, f, dt


def kink_tube(nx=64, ny=None, nz=None, beta=0.05, eta=1.0e-3, kappa_0=0.1, amp=1.0e-3, half_width=4.0):
    """Kink-unstable twisted flux tube (config 5): the uniform-twist field
    B_theta = 2r/(1 + r^2) about the z axis with B_z = 1, from the edge
    vector potential A_z = -ln(1 + r^2) (so the face field is discretely
    divergence-free), rho = 1 and the gas pressure that balances it,
    p = beta/2 + 2/(1 + r^2)^2 (beta = 0.05 in the field B = 1 far from the
    axis; the stated B_z = 1 with this B_theta is not force-free, so the
    pressure rises to 2 + beta/2 on the axis).  Resistive (uniform eta) with
    Braginskii conduction.  The tube is seeded with a helical m = 1 velocity
    v = amp exp(-r^2/2) (cos kz, sin kz, 0), k = 2 pi / Lz.
    x, y in [-half_width, half_width] with walls (BC_USER); z periodic over
    [-Lz/2, Lz/2], Lz = 2 half_width nz / nx (cubic cells).
    Profiles are built in the x-y plane and broadcast along z, so the
    512 x 512 x 1024 case needs no 3-D temporaries."""
    ny = ny or nx
    nz = nz or 2 * nx
    L = half_width
    Lz = 2.0 * L * nz / nx
    grid = Grid(nx, ny, nz, ((-L, L), (-L, L), (-0.5 * Lz, 0.5 * Lz)),
                (BC_USER, BC_USER, BC_USER, BC_USER, BC_PERIODIC, BC_PERIODIC))
    params = default_params(resistive_mhd=1, eta_background=eta, eta0=0.0, j_max=1.0e10,
                            conduction=1, kappa_0=kappa_0, flux_limiter=0.1, heat_flux_limiter=1)
    f = Fields(nx, ny, nz)
    gam = params.gamma
    xb, yb, zb = grid.xb
    xc, yc, zc = grid.xc
    dxb, dyb, dzb = grid.dxb
    # A_z on z-edges (xb_i, yb_j), i, j = -2..n+2
    az = -np.log1p(xb[:, None] ** 2 + yb[None, :] ** 2)
    # bx(i, j) = (Az(i, j) - Az(i, j-1)) / dyb(j): i = -2..nx+2, j = -1..ny+2
    bx2 = (az[:, 1:] - az[:, :-1]) / dyb[None, :]
    # by(i, j) = -(Az(i, j) - Az(i-1, j)) / dxb(i): i = -1..nx+2, j = -2..ny+2
    by2 = -(az[1:, :] - az[:-1, :]) / dxb[:, None]
    f.bx[...] = bx2[:, :, None]
    f.by[...] = by2[:, :, None]
    f.bz[...] = 1.0
    r2 = xc[:, None] ** 2 + yc[None, :] ** 2
    p2 = 0.5 * beta + 2.0 / (1.0 + r2) ** 2
    f.rho[...] = 1.0
    f.energy[...] = (p2 / (gam - 1.0))[:, :, None]
    k = 2.0 * math.pi / Lz
    env = amp * np.exp(-0.5 * (xb[:, None] ** 2 + yb[None, :] ** 2))
    f.vx[...] = env[:, :, None] * np.cos(k * zb)[None, None, :]
    f.vy[...] = env[:, :, None] * np.sin(k * zb)[None, None, :]
    f.vz[...] = 0.0
    return grid, params, f


def partially_ionised(cfg, eos_number, cowling=0, temp_norm=4.0e4, rho_norm=1.0e-6, L_norm=10.0, mf=1.2):
    """A configuration switched to a partially ionised equation of state
    (EOS_ION or EOS_PI, control.f90:187-199), optionally with Cowling
    resistivity.  The normalisation puts the temperatures of the benchmark
    states (T ~ 0.15-0.5 in code units) across the ionisation transition:
    B_norm follows from the requested normalising temperature
    temp_norm = (mbar / k_B) B_norm^2 / (mu0 rho_norm) (normalise.f90:36-40);
    the short L_norm makes eta_perp comparable with the resistive eta of
    the sheet."""
    from .neutral import KB_SI, MH_SI, MU0_SI, set_neutral
    b_norm = math.sqrt(temp_norm * KB_SI / (mf * MH_SI) * MU0_SI * rho_norm)
    set_neutral(cfg[1], b_norm, L_norm, rho_norm, mf, eos_number, cowling)
    return cfg


def uniform(n=8, bcs=(BC_PERIODIC,) * 6, **kw):
    """A uniform state, for invariance tests."""
    grid = Grid(n, n, n, ((0.0, 1.0),) * 3, bcs)
    params = default_params(**kw)
    f = Fields(n, n, n)
    f.rho[...] = 1.3
    f.energy[...] = 0.7
    f.vx[...] = 0.11
    f.vy[...] = -0.07
    f.vz[...] = 0.05
    f.bx[...] = 0.3
    f.by[...] = -0.2
    f.bz[...] = 0.1
    return grid, params, f
