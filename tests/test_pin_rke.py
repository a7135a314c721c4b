"""Pins of the remap kinetic-energy correction, rke (xremap.f90:664-695,
yremap.f90:665-696, zremap.f90:664-695; diagnostics.F90:613-629).

The reference's edge term at the edge between nodes s and s+1 of a sweep,

    ai  = (v_s - f_{s-1}) dm_{s-1} / m_s - (v_s - f_s) dm_s / m_s
    aip = (v_{s+1} - f_s) dm_s / m_{s+1} - (v_{s+1} - f_{s+1}) dm_{s+1} / m_{s+1}
    dk  = [(v_{s+1} - v_s)(f_s - (v_{s+1} + v_s) / 2)
           - ai (v_s - f_s) / 2 + aip (v_{s+1} - f_s) / 2] dm_s / 2,

with m the node masses after the sweep (rho_v1 cv2), f the flux velocity
and dm the vertex mass flux, sums along a periodic line of nodes to exactly
half the change of the line's kinetic energy over the sweep,
sum_s dk = (sum M' v'^2 / 2 - sum M v^2 / 2) / 2, where M = M' - dm_{s-1}
+ dm_s are the masses before the sweep and M' v' = M v + f_{s-1} dm_{s-1}
- f_s dm_s the remapped momenta (the identity is checked below in exact
rational arithmetic).  Each dk is added to the four cells around its edge,
so on a periodic grid the cells' delta_ke sums to 2 (K' - K) over a remap:
the node masses of consecutive sweeps agree, so the sweeps telescope.  A
mirrored state must give the mirrored delta_ke, which fixes where each dk
lands; energy_correction adds the losses (delta_ke < 0) as heat.
"""
from fractions import Fraction
import random

import numpy as np
import pytest

from lare3d_amd.configs import default_params
from lare3d_amd.fields import Fields
from tests.physics import along
from tests.test_pin_oracle3d import _node_mass, _random_periodic


def _edge_terms(M, v, dm, f):
    """The reference's dk on a periodic line (any number type)."""
    n = len(v)
    Mn = [M[i] + dm[i - 1] - dm[i] for i in range(n)]
    out = []
    for s in range(n):
        p, m = (s + 1) % n, s - 1
        ai = (v[s] - f[m]) * dm[m] / Mn[s] - (v[s] - f[s]) * dm[s] / Mn[s]
        aip = (v[p] - f[s]) * dm[s] / Mn[p] - (v[p] - f[p]) * dm[p] / Mn[p]
        dk = (v[p] - v[s]) * (f[s] - (v[p] + v[s]) / 2) - ai * (v[s] - f[s]) / 2 + aip * (v[p] - f[s]) / 2
        out.append(dk * dm[s] / 2)
    return out, Mn


def test_edge_terms_sum_to_half_the_kinetic_energy_change():
    rnd = random.Random(7)
    q = lambda lo, hi: Fraction(rnd.randint(int(lo * 1000), int(hi * 1000)), 1000)
    for n in (3, 5, 8):
        M = [q(1, 3) for _ in range(n)]
        v = [q(-2, 2) for _ in range(n)]
        dm = [q(-0.2, 0.2) for _ in range(n)]
        f = [q(-2, 2) for _ in range(n)]
        dk, Mn = _edge_terms(M, v, dm, f)
        vn = [(M[i] * v[i] + f[i - 1] * dm[i - 1] - f[i] * dm[i]) / Mn[i] for i in range(n)]
        dK = sum(Mn[i] * vn[i] ** 2 for i in range(n)) / 2 - sum(M[i] * v[i] ** 2 for i in range(n)) / 2
        assert sum(dk) == dK / 2


def _kinetic(f, rho_mass, n):
    m = _node_mass(rho_mass, n, 1.0 / n)
    return 0.5 * float(np.sum(m * (f.interior("vx") ** 2 + f.interior("vy") ** 2 + f.interior("vz") ** 2)))


def _rke_state(step, mirror_axis=None, separable=False):
    """A random periodic state (B = 0) through one Lagrangian step and one
    remap with rke on: (fields before the step, after it, after the remap,
    delta_ke).  separable: each velocity component varies along its own
    axis only, gamma = 1 (no pressure) and no viscosity, so that no force
    acts and every sweep's widths (dxb1 of xremap.f90:57-105 and its
    analogues) vary along the sweep axis only."""
    from oracle.oracle import Oracle
    n = 8
    grid, f = _random_periodic(n, seed=31, amp_v=0.4)
    if separable:
        rng = np.random.default_rng(32)
        for a, c in enumerate(("vx", "vy", "vz")):
            getattr(f, c)[...] = along(f.vx.shape, a, 0.4 * rng.standard_normal(f.vx.shape[a]))
        params = default_params(visc1=0.0, visc2=0.0, gamma=1.0)
    else:
        params = default_params(visc1=0.1, visc2=0.5)
    params.rke = 1
    if mirror_axis is not None:
        # the ghosts and the duplicate periodic nodes made consistent first,
        # so that the reflection maps the state onto an equivalent one
        sim = Oracle(grid, params, f)
        sim.boundary_conditions()
        f = _mirror(sim.fields(), mirror_axis)
        sim.close()
    sim = Oracle(grid, params, f)
    try:
        sim.boundary_conditions()
        f0 = sim.fields()
        sim.set_scalars(dt=2.0e-3)
        sim.lagrangian_step()
        fa = sim.fields()
        sim.eulerian_remap(step)
        fb = sim.fields()
        dke = sim.delta_ke()
    finally:
        sim.close()
    return n, f0, fa, fb, dke


@pytest.mark.parametrize("step", [1, 2, 4])
def test_delta_ke_is_twice_the_remap_kinetic_energy_change(step):
    n, f0, fa, fb, dke = _rke_state(step)
    # node masses: before the remap from the cell masses of the start of the
    # step (the Lagrangian step keeps each cell's mass), after it from the
    # remapped density on the fixed grid
    k_a = _kinetic(fa, f0.rho, n)
    k_b = _kinetic(fb, fb.rho, n)
    total = float(np.sum(dke[2:n + 2, 2:n + 2, 2:n + 2]))
    print(f"K before {k_a:.10e} after {k_b:.10e}: change {k_b - k_a:.6e}; sum delta_ke {total:.6e}")
    assert abs(k_b - k_a) > 1e-4 * k_a
    assert abs(total - 2.0 * (k_b - k_a)) <= 1e-9 * abs(k_b - k_a)


def _mirror(f, axis):
    """The state reflected along `axis`: cell I -> n + 1 - I, node i -> n - i
    (both are a reversal of the padded storage), the velocity component
    along the axis negated; axis -1 copies the state unchanged."""
    g = Fields(*f.n)
    for name in ("rho", "energy", "vx", "vy", "vz", "bx", "by", "bz", "visc3", "eta", "visc_dep", "ohmic_dep"):
        a = getattr(f, name)
        if axis >= 0:
            a = np.flip(a, axis)
            if name == ("vx", "vy", "vz")[axis]:
                a = -a
        getattr(g, name)[...] = a
    return g


@pytest.mark.parametrize("step", [1, 3])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_delta_ke_mirrors_with_the_state(axis, step):
    """A reflected state must give the reflected delta_ke, which fixes the
    four cells each edge term goes to.  The reference's step is mirror
    symmetric only in the separable state: its momentum fluxes take the
    sweep widths of the cell on the lower transverse side of a node
    (dxb1(ixp,iy,iz) at node (ix,iy,iz), xremap.f90:643-645), and its
    p_visc takes one edge of each direction per cell (lagran.F90:729-731)."""
    n, _, _, _, dke = _rke_state(step, mirror_axis=-1, separable=True)
    _, _, _, _, dkm = _rke_state(step, mirror_axis=axis, separable=True)
    inner = (slice(2, n + 2),) * 3
    a = np.flip(dkm, axis)[inner]
    b = dke[inner]
    print(f"mirror {axis}: max |delta_ke| {np.abs(b).max():.3e}, max mismatch {np.abs(a - b).max():.3e}")
    assert np.abs(b).std() > 0.1 * np.abs(b).mean()      # delta_ke varies from cell to cell
    assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()


def test_energy_correction_heats_by_the_losses():
    from oracle.oracle import Oracle
    n = 8
    grid, f = _random_periodic(n, seed=31, amp_v=0.4)
    params = default_params(visc1=0.1, visc2=0.5)
    params.rke = 1
    sim = Oracle(grid, params, f)
    try:
        sim.boundary_conditions()
        sim.set_scalars(dt=2.0e-3)
        sim.lagrangian_step()
        sim.eulerian_remap(2)
        fb = sim.fields()
        dke = sim.delta_ke()[2:n + 2, 2:n + 2, 2:n + 2]
        sim.energy_correction()
        fc = sim.fields()
    finally:
        sim.close()
    cv = (1.0 / n) ** 3
    heat = fb.interior("rho") * cv * (fc.interior("energy") - fb.interior("energy"))
    loss = np.maximum(-dke, 0.0)
    assert (dke > 0).any() and (dke < 0).any()        # both signs occur, so the clipping matters
    assert np.abs(heat - loss).max() <= 1e-12 * loss.max()
    assert np.array_equal(fc.interior("rho"), fb.interior("rho"))
