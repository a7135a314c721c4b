"""Inputs of the device initial-condition routines.

potential_inputs() builds the argument of lare_dev_potential_field
(initial_conditions.f90:114-296): the SOR parameters, the inverse squared
spacings (n_global / length)^2 of :126-128, and the flux patch the lower z
boundary adds to phi each iteration (:285-293),
    phi(ix, iy, 0) = phi(ix, iy, 1) + amp dzc(1) (1 - tanh((r - radius) / width)),
    r = sqrt((xc(ix) - centre)^2 + yc(iy)^2),
evaluated here on the host so that the device and the CPU restatement add
the same bits.  The reference's example uses centre = 10, radius = 1,
amp = 1, width = 0.2, w = 1.6 and a tolerance of 5e-5.
"""
import numpy as np

from .abi import LarePotential, dptr


def potential_inputs(grid, centre=10.0, radius=1.0, amp=1.0, width=0.2, omega=1.6, tol=5.0e-5,
                     max_iterations=1000000):
    nx, ny, nz = grid.n
    xc = grid.xc[0][2:nx + 2]  # xc(1:nx)
    yc = grid.xc[1][2:ny + 2]
    dzc1 = grid.dxc[2][2]      # dzc(1)
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    r = np.sqrt((X - centre) ** 2 + Y ** 2)
    flux0 = np.asfortranarray(amp * dzc1 * (1.0 - np.tanh((r - radius) / width)))
    pf = LarePotential()
    pf.omega = omega
    pf.fractional_error = tol
    d = []
    for a in range(3):
        length = grid.extent[a][1] - grid.extent[a][0]
        q = grid.n[a] / length
        d.append(q * q)
    pf.dx1, pf.dy1, pf.dz1 = d
    flux0 = flux0.ravel(order="F")
    pf.flux0 = dptr(flux0)
    pf.max_iterations = max_iterations
    pf._keep = flux0  # the pointer above refers to this array
    return pf
