"""Closed-form references used to pin the CPU restatement (and the device).

Nothing here restates a reference routine: each function is an independent
textbook result (exact Riemann solver, Chebyshev stability polynomial of
RKL2, linear wave speeds, discrete Laplacian eigenvalues).
"""
import math

import numpy as np

from lare3d_amd import BC_PERIODIC, BC_USER, Fields, Grid
from lare3d_amd.configs import default_params


def exact_riemann(rhoL, uL, pL, rhoR, uR, pR, gamma, x, t, x0):
    """Exact solution of the ideal-gas Riemann problem (Toro, Riemann Solvers
    and Numerical Methods for Fluid Dynamics, ch. 4) sampled at x."""
    cL, cR = math.sqrt(gamma * pL / rhoL), math.sqrt(gamma * pR / rhoR)
    g1 = (gamma - 1.0) / (2.0 * gamma)
    g2 = (gamma + 1.0) / (2.0 * gamma)

    def f(p, rho, pk, c):
        if p > pk:  # shock
            A = 2.0 / ((gamma + 1.0) * rho)
            B = (gamma - 1.0) / (gamma + 1.0) * pk
            return (p - pk) * math.sqrt(A / (p + B)), math.sqrt(A / (B + p)) * (1.0 - 0.5 * (p - pk) / (B + p))
        pr = p / pk  # rarefaction
        return 2.0 * c / (gamma - 1.0) * (pr ** g1 - 1.0), 1.0 / (rho * c) * pr ** (-g2)

    p = 0.5 * (pL + pR)
    for _ in range(100):
        fL, dL = f(p, rhoL, pL, cL)
        fR, dR = f(p, rhoR, pR, cR)
        dp = (fL + fR + uR - uL) / (dL + dR)
        p = max(p - dp, 1e-12)
        if abs(dp) < 1e-14 * p:
            break
    fL, _ = f(p, rhoL, pL, cL)
    fR, _ = f(p, rhoR, pR, cR)
    u = 0.5 * (uL + uR) + 0.5 * (fR - fL)
    out = np.empty((len(x), 3))
    for q, xx in enumerate(x):
        s = (xx - x0) / t
        if s <= u:  # left of contact
            if p > pL:
                sL = uL - cL * math.sqrt(g2 * p / pL + g1)
                if s < sL:
                    out[q] = (rhoL, uL, pL)
                else:
                    r = rhoL * ((p / pL + (gamma - 1) / (gamma + 1)) / ((gamma - 1) / (gamma + 1) * p / pL + 1))
                    out[q] = (r, u, p)
            else:
                cs = cL * (p / pL) ** g1
                sh, st = uL - cL, u - cs
                if s < sh:
                    out[q] = (rhoL, uL, pL)
                elif s > st:
                    out[q] = (rhoL * (p / pL) ** (1 / gamma), u, p)
                else:
                    c = 2 / (gamma + 1) * (cL + (gamma - 1) / 2 * (uL - s))
                    uu = 2 / (gamma + 1) * (cL + (gamma - 1) / 2 * uL + s)
                    r = rhoL * (c / cL) ** (2 / (gamma - 1))
                    out[q] = (r, uu, pL * (c / cL) ** (2 * gamma / (gamma - 1)))
        else:
            if p > pR:
                sR = uR + cR * math.sqrt(g2 * p / pR + g1)
                if s > sR:
                    out[q] = (rhoR, uR, pR)
                else:
                    r = rhoR * ((p / pR + (gamma - 1) / (gamma + 1)) / ((gamma - 1) / (gamma + 1) * p / pR + 1))
                    out[q] = (r, u, p)
            else:
                cs = cR * (p / pR) ** g1
                sh, st = uR + cR, u + cs
                if s > sh:
                    out[q] = (rhoR, uR, pR)
                elif s < st:
                    out[q] = (rhoR * (p / pR) ** (1 / gamma), u, p)
                else:
                    c = 2 / (gamma + 1) * (cR - (gamma - 1) / 2 * (uR - s))
                    uu = 2 / (gamma + 1) * (-cR + (gamma - 1) / 2 * uR + s)
                    r = rhoR * (c / cR) ** (2 / (gamma - 1))
                    out[q] = (r, uu, pR * (c / cR) ** (2 * gamma / (gamma - 1)))
    return out


def legendre_P(s, x):
    """Legendre polynomial P_s(x) by Bonnet's recursion."""
    p0, p1 = 1.0, x
    if s == 0:
        return p0
    for n in range(1, s):
        p0, p1 = p1, ((2 * n + 1) * x * p1 - n * p0) / (n + 1)
    return p1


def rkl2_amplification(s, z):
    """Stability polynomial of the s-stage RKL2 scheme for y' = (z / dt) y
    (Meyer, Balsara & Aslam 2012, MNRAS 422, 2102): R_s(z) = a_s + b_s
    P_s(w0 + w1 z) with Legendre P_s, w0 = 1, w1 = 4 / (s^2 + s - 2),
    b_s = (s^2 + s - 2) / (2 s (s + 1)), a_s = 1 - b_s.  R_s(z) agrees with
    exp(z) to second order (P_s'(1) = s(s+1)/2)."""
    w1 = 4.0 / (s * s + s - 2.0)
    b = (s * s + s - 2.0) / (2.0 * s * (s + 1.0))
    return (1.0 - b) + b * legendre_P(s, 1.0 + w1 * z)


def laplacian_eig(k, dx):
    """Eigenvalue magnitude of the 3-point Laplacian for a mode of wavenumber k."""
    return 4.0 / dx ** 2 * math.sin(0.5 * k * dx) ** 2


def line_problem(n, axis, ntrans=4, extent=(0.0, 1.0), bc_axis=BC_USER, bc_trans=BC_PERIODIC, **kw):
    """A grid that is n long on `axis` and ntrans on the others, with
    params; returns (grid, params, fields, coordinate arrays)."""
    ns = [ntrans, ntrans, ntrans]
    ns[axis] = n
    ext = [(0.0, ntrans / n * (extent[1] - extent[0]))] * 3
    ext[axis] = extent
    bcs = [bc_trans] * 6
    bcs[2 * axis] = bcs[2 * axis + 1] = bc_axis
    grid = Grid(ns[0], ns[1], ns[2], tuple(ext), tuple(bcs))
    params = default_params(**kw)
    return grid, params, Fields(*ns)


def along(arr_shape, axis, values):
    """Broadcast a 1-D profile along `axis` over an array shape."""
    sh = [1, 1, 1]
    sh[axis] = len(values)
    return np.broadcast_to(np.asarray(values).reshape(sh), arr_shape)
