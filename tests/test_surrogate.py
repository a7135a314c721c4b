"""Learned conduction surrogate (no counterpart in the reference).

CPU tests pin the numpy restatement (oracle/surrogate_ref.py): BF16 rounding
against bit patterns the format fixes, the backward pass against central
finite differences, the features against closed forms.  GPU tests compare
the device kernels with it through the C-ABI: forward and gradient of the
BF16 MFMA network, the training rows drawn from a state, and the RKL2
targets; then train on seeded states and show what the trained network gains
over the untrained one and over the plain solver, with the fallback.
"""
import math
import time

import numpy as np
import pytest

from lare3d_amd import configs
from oracle import surrogate_ref as ref

# --------------------------------------------------------------------------
# CPU: pins of the restatement
# --------------------------------------------------------------------------


def _bits(x):
    return int(np.asarray([x], dtype=np.float32).view(np.uint32)[0])


def _f32(bits):
    return np.asarray([bits], dtype=np.uint32).view(np.float32)[0]


def test_bf16_rounding_bit_patterns():
    """BF16 keeps the top 16 bits of an FP32 word; round to nearest, ties to
    the even upper half."""
    one = 1.0
    cases = [
        (one, 0x3F800000),                              # exact
        (one + 2.0 ** -8, 0x3F800000),                  # tie -> even (down)
        (one + 3 * 2.0 ** -8, 0x3F820000),              # tie -> even (up)
        (one + 2.0 ** -8 + 2.0 ** -20, 0x3F810000),     # above the tie -> up
        (-(one + 2.0 ** -8 + 2.0 ** -20), 0xBF810000),  # sign symmetric
        (_f32(0x7F61B1E6), 0x7F620000),                 # low half 0xB1E6 > 0x8000 -> up
        (_f32(0x7F617FFF), 0x7F610000),                 # low half below the tie -> down
        (_f32(0x000116C2), 0x00010000),                 # subnormal, down
        (_f32(0x00018000), 0x00020000),                 # subnormal tie, odd -> up
        (_f32(0x00028000), 0x00020000),                 # subnormal tie, even -> stays
        (_f32(0x7F7FFFFF), 0x7F800000),                 # largest finite rounds to infinity
    ]
    for x, want in cases:
        got = _bits(ref.bf16(np.float32(x)))
        assert got == want, (x, hex(got), hex(want))


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(7)
    w = ref.init_params(3).astype(np.float64)
    w[ref.OFF_B1:ref.OFF_W2] = 0.1 * rng.standard_normal(ref.H)
    w[ref.OFF_B3] = 0.05
    X = rng.standard_normal((6, ref.K))
    X[:, ref.KREAL:] = 0.0
    t = rng.standard_normal(6)
    g, loss = ref.backward(w, X, t, emulate_bf16=False)
    e = ref.forward(w, X, emulate_bf16=False) - t
    assert abs(loss - np.mean(e * e)) <= 1e-14 * loss
    blocks = [(ref.OFF_W1, ref.OFF_B1), (ref.OFF_B1, ref.OFF_W2), (ref.OFF_W2, ref.OFF_B2),
              (ref.OFF_B2, ref.OFF_W3), (ref.OFF_W3, ref.OFF_B3), (ref.OFF_B3, ref.NPARAM)]
    h = 1e-6
    for a, b in blocks:
        idx = rng.choice(np.arange(a, b), size=min(8, b - a), replace=False)
        if a == ref.OFF_W1:  # only rows that meet non-zero inputs carry gradient
            idx = [i for i in idx if (i - a) // ref.H < ref.KREAL] or [a]
        for i in idx:
            wp, wm = w.copy(), w.copy()
            wp[i] += h
            wm[i] -= h
            lp = np.mean((ref.forward(wp, X, emulate_bf16=False) - t) ** 2)
            lm = np.mean((ref.forward(wm, X, emulate_bf16=False) - t) ** 2)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * (1.0 + abs(fd)), (a, i, fd, g[i])


def test_features_closed_form_on_a_linear_temperature():
    # uniform B, rho; energy linear in x: T = tfac e gives dT/dx = tfac * slope
    grid, params, f = configs.uniform(8)
    slope = 0.3
    xc = grid.xc[0]
    for i in range(-1, 8 + 3):
        f.energy[i + 1, :, :] = 2.0 + slope * xc[i + 1]
    dt = 0.0123
    tfac = (params.gamma - 1.0) / (2.0 - params.xi_n)
    cells = [(1, 1, 1), (4, 5, 6), (8, 8, 8)]
    X = ref.features(grid, f, dt, params.gamma, params.xi_n, cells)
    bm = math.sqrt(0.3 ** 2 + 0.2 ** 2 + 0.1 ** 2)
    for r, (i, j, k) in enumerate(cells):
        row = X[r].reshape(-1)[: ref.KREAL].reshape(27, ref.CH)
        centre = row[13]
        want = [tfac * (2.0 + slope * xc[i + 1]), 0.3 / bm, -0.2 / bm, 0.1 / bm, 1.3, math.log(dt), bm,
                tfac * slope, 0.0, 0.0, 0.0, 0.0]
        np.testing.assert_allclose(centre, np.float32(want), rtol=2e-6, atol=1e-7)
        # neighbour (dx = +1) is the centre of cell i+1
        assert row[14][0] == np.float32(tfac * (2.0 + slope * xc[i + 2]))
    assert np.all(X[:, ref.KREAL:] == 0.0)


def test_sample_cells_are_interior_seeded_and_biased():
    n = (16, 12, 20)
    # a temperature with one hot cell: every odd row whose candidates include
    # the hot cell or a face neighbour of it picks such a cell
    hot = (5, 6, 7)
    T = lambda i, j, k: np.float32(10.0 if (i, j, k) == hot else 1.0)
    a = ref.sample_cells(5, 500, n, T)
    b = ref.sample_cells(5, 500, n, T)
    c = ref.sample_cells(6, 500, n, T)
    assert a == b and a != c
    for i, j, k in a:
        assert 1 <= i <= n[0] and 1 <= j <= n[1] and 1 <= k <= n[2]
    # every interior index along each axis is reachable by the uniform rows
    even = a[0::2]
    assert len({i for i, _, _ in even}) == n[0] and len({k for _, _, k in even}) == n[2]
    # odd rows prefer the active region: the hot cell and its 6 neighbours
    near = {hot} | {(hot[0] + d[0], hot[1] + d[1], hot[2] + d[2]) for d in
                    ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))}
    frac_odd = np.mean([p in near for p in a[1::2]])
    frac_even = np.mean([p in near for p in even])
    assert frac_odd > 4 * frac_even


# --------------------------------------------------------------------------
# GPU: device kernels against the restatement
# --------------------------------------------------------------------------


def _device(grid, params, f):
    from lare3d_amd.device import Device
    return Device(grid, params, f)


@pytest.mark.gpu
def test_device_forward_and_gradient_match_restatement():
    from lare3d_amd.surrogate import Surrogate
    grid, params, f, dt = configs.harris_conduction(16, seed=1)
    dev = _device(grid, params, f)
    try:
        sur = Surrogate(dev)
        rng = np.random.default_rng(11)
        w = ref.init_params(5)
        w[ref.OFF_B1:ref.OFF_W2] = (0.1 * rng.standard_normal(ref.H)).astype(np.float32)
        w[ref.OFF_B2:ref.OFF_W3] = (0.1 * rng.standard_normal(ref.H)).astype(np.float32)
        w[ref.OFF_B3] = 0.03
        sur.set_params(w)
        for rows in (100, 512):  # 100 is padded to 112 columns on the device
            print(f"\n[case] surrogate forward/grad rows={rows}", flush=True)
            X = rng.standard_normal((rows, ref.K)).astype(np.float32)
            X[:, ref.KREAL:] = 0.0
            t = rng.standard_normal(rows).astype(np.float32)
            y_dev = sur.forward(X)
            y_orc = ref.forward(w, X, emulate_bf16=True)
            err = np.max(np.abs(y_dev - y_orc)) / np.max(np.abs(y_orc))
            print(f"forward max rel err {err:.2e}")
            assert err <= 2e-3
            g_dev, loss_dev = sur.grad(X, t)
            g_orc, loss_orc = ref.backward(w, X, t, emulate_bf16=True)
            assert abs(loss_dev - loss_orc) <= 2e-3 * loss_orc
            for name, a, b in (("W1", ref.OFF_W1, ref.OFF_B1), ("b1", ref.OFF_B1, ref.OFF_W2),
                               ("W2", ref.OFF_W2, ref.OFF_B2), ("b2", ref.OFF_B2, ref.OFF_W3),
                               ("W3", ref.OFF_W3, ref.OFF_B3), ("b3", ref.OFF_B3, ref.NPARAM)):
                rel = np.linalg.norm(g_dev[a:b] - g_orc[a:b]) / np.linalg.norm(g_orc[a:b])
                print(f"grad {name}: rel L2 err {rel:.2e}")
                assert rel <= 2e-2, name
    finally:
        dev.close()


@pytest.mark.gpu
def test_device_inference_matches_restatement():
    """One surrogate super-step accepted unconditionally (tolerance 1e30) leaves
    E1 = E0 + dT / tfac on every interior cell, dT being the network's output
    times the target scale: compare with the restatement's forward pass (BF16
    operands) on the standardised features of the same state.  The 16^3 grid
    leaves the inference tiles partial in x (64-cell rows)."""
    from lare3d_amd.surrogate import Surrogate
    grid, params, f, dt = configs.harris_conduction(16, seed=3, s_target=5)
    dev = _device(grid, params, f)
    try:
        dev.boundary_conditions()
        dev.set_scalars(dt=dt)
        f0 = dev.fields()
        sur = Surrogate(dev)
        w = ref.init_params(7)
        w[ref.OFF_B3] = 0.02
        sur.set_params(w)
        print("\n[case] surrogate inference 16^3", flush=True)
        sur.collect(2000, seed=5)
        sur.fit_norm()
        mu, sd, y_scale = sur.norm()
        sur.enable(False, 1.0e30)
        res, acc = sur.conduct()
        assert acc
        f1 = dev.fields()
    finally:
        dev.close()
    n = grid.nx
    cells = [(i, j, k) for k in range(1, n + 1) for j in range(1, n + 1) for i in range(1, n + 1)]
    X = ref.features(grid, f0, dt, params.gamma, params.xi_n, cells)
    for q in range(27):
        for ch in range(ref.NCH):
            col = q * ref.CH + ch
            X[:, col] = (X[:, col] - mu[ch]) / sd[ch]
    y = ref.forward(w, X, emulate_bf16=True)
    dT_orc = (y * np.float32(y_scale)).astype(np.float64)
    tfac = (params.gamma - 1.0) / (2.0 - params.xi_n)
    e0 = np.array([f0.energy[i + 1, j + 1, k + 1] for i, j, k in cells])
    e1 = np.array([f1.energy[i + 1, j + 1, k + 1] for i, j, k in cells])
    dT_dev = tfac * (e1 - e0)
    err = float(np.linalg.norm(dT_dev - dT_orc) / np.linalg.norm(dT_orc))
    print(f"inference dT rel L2 err {err:.2e} over {len(cells)} cells")
    assert err <= 2e-3


@pytest.mark.gpu
def test_device_training_rows_match_restatement():
    from oracle.oracle import Oracle
    from lare3d_amd.surrogate import Surrogate
    grid, params, f, dt = configs.harris_conduction(16, seed=2, s_target=5)
    dev = _device(grid, params, f)
    try:
        dev.boundary_conditions()
        dev.set_scalars(dt=dt)
        f_dev = dev.fields()
        sur = Surrogate(dev)
        rows, seed = 96, 77
        print("\n[case] surrogate collect 16^3 rows=96", flush=True)
        assert sur.collect(rows, seed) == rows
        X, y = sur.rows(rows)
        tfac = (params.gamma - 1.0) / (2.0 - params.xi_n)
        temp = lambda i, j, k: np.float32(tfac * f_dev.energy[i + 1, j + 1, k + 1])
        cells = ref.sample_cells(seed, rows, grid.n, temp)
        Xo = ref.features(grid, f_dev, dt, params.gamma, params.xi_n, cells)
        np.testing.assert_allclose(X, Xo, rtol=1e-6, atol=1e-6)
        # targets: one plain super-step of the CPU restatement from the same state
        orc = Oracle(grid, params, f_dev)
        orc.set_scalars(dt=dt)
        orc.conduct_heat()
        f1 = orc.fields()
        orc.close()
        yo = np.array([tfac * (f1.energy[i + 1, j + 1, k + 1] - f_dev.energy[i + 1, j + 1, k + 1])
                       for i, j, k in cells])
        err = np.max(np.abs(y - yo)) / np.max(np.abs(yo))
        print(f"targets max rel err {err:.2e}, max |dT| {np.max(np.abs(yo)):.3e}")
        assert err <= 1e-5
        # the state is unchanged by the collection
        f_after = dev.fields()
        assert np.array_equal(f_after.energy, f_dev.energy)
    finally:
        dev.close()


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.gpu
def test_training_gain_and_fallback():
    """Train on 50k rows from seeded Harris + conduction states (32^3), then on
    a held-out state compare the predicted super-step with the plain RKL2
    one: the trained network must beat the untrained one; the residual check
    accepts or falls back, and a fallback reproduces RKL2 exactly."""
    from lare3d_amd.surrogate import Surrogate, initial_params
    n = 32
    grid, params, f, dt = configs.harris_conduction(n, seed=0, s_target=3)
    dev = _device(grid, params, f)
    try:
        sur = Surrogate(dev)
        sur.set_params(initial_params(0))
        total = 0
        t0 = time.time()
        for s in range(10):
            g_, p_, f_, dt_ = configs.harris_conduction(n, seed=100 + s, s_target=3 + (s % 3) * 2)
            dev.upload(f_)
            dev.boundary_conditions()
            dev.set_scalars(dt=dt_)
            total = sur.collect(5000, seed=500 + s)
        print(f"\n[case] collected {total} rows in {time.time() - t0:.2f} s", flush=True)
        assert total == 50000
        sur.fit_norm()
        w0 = sur.params()
        t0 = time.time()
        # Adam with a stepped learning rate (the moments carry over between calls)
        loss = np.concatenate([sur.train(2000, batch=512, lr=lr, seed=9 + q)
                               for q, lr in enumerate((1e-3, 3e-4, 1e-4))])
        print(f"trained {loss.size} steps in {time.time() - t0:.2f} s: loss {loss[:50].mean():.3e} -> "
              f"{loss[-50:].mean():.3e}")
        assert loss[-50:].mean() < 0.2 * loss[:50].mean()

        # held-out state
        g_, p_, fh, dth = configs.harris_conduction(n, seed=999, s_target=5)
        tfac = (params.gamma - 1.0) / (2.0 - params.xi_n)

        def plain_dT():
            dev.upload(fh)
            dev.boundary_conditions()
            dev.set_scalars(dt=dth)
            dev.conduct_heat()
            return dev.fields()

        f_rkl2 = plain_dT()
        dT_ref = tfac * (f_rkl2.interior("energy") - fh.interior("energy"))

        def sur_dT(tol):
            dev.upload(fh)
            dev.boundary_conditions()
            dev.set_scalars(dt=dth)
            sur.enable(False, tol)
            res, acc = sur.conduct()
            return res, acc, dev.fields()

        res_t, acc_t, f_t = sur_dT(1e30)          # force acceptance: the network's own answer
        dT_t = tfac * (f_t.interior("energy") - fh.interior("energy"))
        err_trained = _rel_l2(dT_t, dT_ref)
        # the same with the untrained parameters (same normalisation)
        w_trained = sur.params()
        sur.set_params(w0)
        res_u, acc_u, f_u = sur_dT(1e30)
        dT_u = tfac * (f_u.interior("energy") - fh.interior("energy"))
        err_untrained = _rel_l2(dT_u, dT_ref)
        sur.set_params(w_trained)
        T_err = _rel_l2(tfac * f_t.interior("energy"), tfac * f_rkl2.interior("energy"))
        print(f"held-out dT rel L2: trained {err_trained:.3e}, untrained {err_untrained:.3e}; "
              f"T rel L2 {T_err:.3e}; residual trained {res_t:.3e}, untrained {res_u:.3e}")
        assert err_trained < 0.5 * err_untrained
        assert res_t < res_u

        dev.upload(fh)
        dev.boundary_conditions()
        dev.set_scalars(dt=dth)
        e_before = dev.fields().energy
        r_plain = sur.plain_residual()
        print(f"residual check applied to the plain RKL2 step: {r_plain:.3e}")
        assert r_plain < res_t
        assert np.array_equal(dev.fields().energy, e_before)  # state unchanged
        # tolerance below the residual: the step falls back to RKL2 exactly
        res_f, acc_f, f_f = sur_dT(0.5 * res_t)
        assert not acc_f
        assert np.array_equal(f_f.energy, f_rkl2.energy)
        a, rj, last = sur.stats()
        print(f"stats: accepted {a}, rejected {rj}, last residual {last:.3e}")
        assert rj >= 1 and a >= 2
    finally:
        dev.close()
