"""Pins of potential_field (initial_conditions.f90:114-296), the device's
initial-condition generator, on the CPU restatement; the device runs the
same inputs through the C-ABI and must give the same bits.

* One SOR iteration from phi = 0 has a closed form.  The first colour pass
  ((ix + iy) even, z1 = 1) sees phi = 0 everywhere and changes nothing;
  phi_mpi then puts the flux patch s(ix, iy) into phi(ix, iy, 0).  In the
  second pass each odd column is relaxed from iz = 1 upwards with the x/y
  neighbours still zero and the point below already updated (the colour
  does not depend on iz, so a column is a Gauss-Seidel sweep), so
      phi(ix, iy, k) = s(ix, iy) a^k,   a = w dz1 / (2 (dx1 + dy1 + dz1)),
  and the largest residual is max(s) dz1 / (2 (dx1 + dy1 + dz1)).  A
  colouring in all three dimensions, a missing w, another normalisation or
  a source on the wrong boundary all change this.
* At convergence (tolerance 1e-13) phi solves the discrete Laplace problem
  with the reference's boundary rules: zero-gradient x and y walls
  (phi(0) = phi(1), phi(n+1) = phi(n)), the flux patch below
  (phi(0) = phi(1) + s), phi = 0 above.  The same problem is assembled here
  as a sparse matrix and solved directly; phi and the scaled face field
  agree to round-off of the solve.
* The field is B = -grad phi, scaled by the largest bz on the lower
  boundary: that bz is 1, and div B vanishes to the tolerance.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from lare3d_amd import BC_PERIODIC, BC_USER
from lare3d_amd.configs import default_params
from lare3d_amd.fields import Fields
from lare3d_amd.grid import Grid
from lare3d_amd.initial import potential_inputs
from tests.backends import BACKENDS, close, make


def _setup(backend, n=(10, 12, 12), ext=((-5.0, 15.0), (-6.0, 6.0), (0.0, 18.0)), xbc=BC_USER):
    grid = Grid(*n, ext, (xbc, xbc, BC_USER, BC_USER, BC_USER, BC_USER))
    f = Fields(*n)
    f.rho[...] = 1.0
    f.energy[...] = 1.0
    sim = make(backend, grid, default_params(), f)
    return grid, sim


def test_one_iteration_closed_form():
    grid, sim = _setup("oracle")
    pf = potential_inputs(grid, centre=4.0, radius=2.0, max_iterations=1, tol=0.0)
    it, res, bm, phi = sim.potential_field(pf, want_phi=True)
    close(sim)
    nx, ny, nz = grid.n
    s = pf._keep.reshape((nx, ny), order="F")
    tot = pf.dx1 + pf.dy1 + pf.dz1
    a = pf.omega * pf.dz1 / (2.0 * tot)
    assert it == 1
    assert res == pytest.approx(s.max() * pf.dz1 / (2.0 * tot), rel=1e-14)
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            col = phi[ix + 1, iy + 1, 2:nz + 2]  # phi(ix, iy, 1:nz)
            if (ix + iy) % 2 == 0:
                assert not np.any(col)
            else:
                want = s[ix - 1, iy - 1] * a ** np.arange(1, nz + 1)
                np.testing.assert_allclose(col, want, rtol=1e-13, atol=0)


def _direct(grid, pf):
    """phi(1:n) of the converged discrete problem, by sparse LU."""
    nx, ny, nz = grid.n
    s = pf._keep.reshape((nx, ny), order="F")
    N = nx * ny * nz
    idx = lambda i, j, k: i + nx * (j + ny * k)
    rows, cols, vals, b = [], [], [], np.zeros(N)
    w = (pf.dx1, pf.dy1, pf.dz1)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = idx(i, j, k)
                rows.append(r); cols.append(r); vals.append(-2.0 * sum(w))
                for ax, (d, lim) in enumerate(((i, nx), (j, ny), (k, nz))):
                    for step in (-1, 1):
                        q = [i, j, k]
                        q[ax] += step
                        if 0 <= q[ax] < lim:
                            rows.append(r); cols.append(idx(*q)); vals.append(w[ax])
                        elif ax < 2:
                            rows.append(r); cols.append(r); vals.append(w[ax])  # zero gradient
                        elif step < 0:
                            rows.append(r); cols.append(r); vals.append(w[ax])  # phi(0) = phi(1) + s
                            b[r] -= w[ax] * s[i, j]
                        # phi(nz + 1) = 0 adds nothing
    A = sp.csr_matrix((vals, (rows, cols)), shape=(N, N))
    return spla.spsolve(A.tocsc(), b).reshape((nx, ny, nz), order="F")


def test_converged_solution_is_the_direct_solve():
    grid, sim = _setup("oracle")
    pf = potential_inputs(grid, centre=4.0, radius=2.0, tol=1e-13, max_iterations=200000)
    it, res, bm, phi = sim.potential_field(pf, want_phi=True)
    f = sim.fields()
    close(sim)
    nx, ny, nz = grid.n
    want = _direct(grid, pf)
    got = phi[2:nx + 2, 2:ny + 2, 2:nz + 2]
    print(f"iterations {it}, residual {res:.2e}, bz_max {bm:.6f}")
    assert res < 1e-13 and it < 200000
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10 * np.abs(want).max())
    # B = -grad phi / bz_max on the faces the reference sets
    dxc, dyc, dzc = grid.dxc
    bx = -(want[1:, :, :] - want[:-1, :, :]) / dxc[3:nx + 2][:, None, None] / bm
    np.testing.assert_allclose(f.view("bx", 1, nx - 1, 1, ny, 1, nz), bx, rtol=1e-9, atol=1e-12)
    assert f.view("bz", 1, nx, 1, min(ny, nz), 0, 0).max() == 1.0
    en_divb = np.abs((f.view("bx", 1, nx, 1, ny, 1, nz) - f.view("bx", 0, nx - 1, 1, ny, 1, nz)) / grid.dxb[0][2:nx + 2][:, None, None]
                     + (f.view("by", 1, nx, 1, ny, 1, nz) - f.view("by", 1, nx, 0, ny - 1, 1, nz)) / grid.dxb[1][2:ny + 2][None, :, None]
                     + (f.view("bz", 1, nx, 1, ny, 1, nz) - f.view("bz", 1, nx, 1, ny, 0, nz - 1)) / grid.dxb[2][2:nz + 2][None, None, :])
    assert en_divb.max() < 1e-9


@pytest.mark.parametrize("xbc", [BC_USER, BC_PERIODIC])
@pytest.mark.parametrize("backend", BACKENDS)
def test_potential_field_runs(backend, xbc):
    """Reference tolerance 5e-5: converges, bz_max scales the lower boundary
    field to 1; the device gives the oracle's bits (iterations, residual,
    every face value)."""
    from oracle.oracle import Oracle
    n, ext = (12, 12, 12), ((-5.0, 15.0), (-10.0, 10.0), (0.0, 20.0))
    grid, sim = _setup(backend, n, ext, xbc)
    pf = potential_inputs(grid)
    it, res, bm = sim.potential_field(pf)
    f = sim.fields()
    close(sim)
    assert res < 5e-5 and 1 < it < 1000000
    assert f.view("bz", 1, 12, 1, 12, 0, 0).max() == 1.0
    if backend == "device":
        g2, orc = _setup("oracle", n, ext, xbc)
        ito, reso, bmo = orc.potential_field(potential_inputs(g2))
        fo = orc.fields()
        close(orc)
        print(f"device {it} {res!r} {bm!r} / oracle {ito} {reso!r} {bmo!r}")
        assert (it, res, bm) == (ito, reso, bmo)
        for name in ("bx", "by", "bz"):
            assert np.array_equal(getattr(f, name), getattr(fo, name)), name
