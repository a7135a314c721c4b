"""Learned conduction surrogate on the device (no counterpart in the reference).

A 3x3x3-stencil MLP stands in for whole RKL2 conduction super-steps
(conduct.f90:39); each use is checked by a trapezoidal residual and falls
back to the plain RKL2 solver above the tolerance.  Training data, training
(forward and backward BF16 MFMA kernels, Adam) and inference all run in the
device library; this module drives them through the C-ABI.

    sur = Surrogate(dev)
    sur.set_params(initial_params(seed))
    for state in training states:            # Device.upload of each state
        sur.collect(rows, seed)
    sur.fit_norm()
    losses = sur.train(steps, batch, lr, seed)
    dev_residual, accepted = sur.conduct()   # one checked super-step
    sur.enable(True, tol=1e-4)               # use it inside lagrangian_step
"""
import ctypes as C
import math

import numpy as np

from .abi import check

K = 336            # input row length (27 cells x 12 channels, padded)
CH = 12
NCH = 10
H = 64
NPARAM = K * H + H + H * H + H + H + 1


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def initial_params(seed=0):
    """Seeded initial parameters: W1 ~ N(0, 1/270) on the rows of the ten used
    channels of the 27 cells (zero elsewhere), W2 ~ N(0, 1/64),
    W3 ~ N(0, 0.01/64), zero biases."""
    rng = np.random.default_rng(seed)
    W1 = np.zeros((K, H))
    used = np.array([(kk < 27 * CH) and (kk % CH < NCH) for kk in range(K)])
    W1[used] = rng.standard_normal((int(used.sum()), H)) / math.sqrt(27 * NCH)
    W2 = rng.standard_normal((H, H)) / math.sqrt(H)
    W3 = 0.1 * rng.standard_normal(H) / math.sqrt(H)
    return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), W3, [0.0]]).astype(np.float32)


class Surrogate:
    """The surrogate of one Device (shares its library state)."""

    def __init__(self, dev):
        self.dev = dev
        self.lib = dev.lib

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        check(self.lib, self.lib.lare_dev_sur_set_params(_fp(w), w.size), "sur_set_params")

    def params(self):
        w = np.zeros(NPARAM, dtype=np.float32)
        check(self.lib, self.lib.lare_dev_sur_get_params(_fp(w), w.size), "sur_get_params")
        return w

    def set_norm(self, mu, sd, y_scale):
        mu = np.ascontiguousarray(mu, dtype=np.float32)
        sd = np.ascontiguousarray(sd, dtype=np.float32)
        check(self.lib, self.lib.lare_dev_sur_set_norm(_fp(mu), _fp(sd), float(y_scale)), "sur_set_norm")

    def norm(self):
        mu = np.zeros(NCH, dtype=np.float32)
        sd = np.zeros(NCH, dtype=np.float32)
        ys = C.c_float()
        check(self.lib, self.lib.lare_dev_sur_get_norm(_fp(mu), _fp(sd), C.byref(ys)), "sur_get_norm")
        return mu, sd, ys.value

    def collect(self, rows, seed):
        total = C.c_int32()
        check(self.lib, self.lib.lare_dev_sur_collect(int(rows), int(seed), C.byref(total)), "sur_collect")
        return total.value

    def rows(self, count):
        X = np.zeros((count, K), dtype=np.float32)
        y = np.zeros(count, dtype=np.float32)
        check(self.lib, self.lib.lare_dev_sur_rows(_fp(X), _fp(y), int(count)), "sur_rows")
        return X, y

    def fit_norm(self):
        check(self.lib, self.lib.lare_dev_sur_fit_norm(), "sur_fit_norm")

    def train(self, steps, batch=512, lr=1e-3, seed=1):
        loss = np.zeros(steps, dtype=np.float32)
        check(self.lib, self.lib.lare_dev_sur_train(int(steps), int(batch), float(lr), int(seed), _fp(loss)),
              "sur_train")
        return loss

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.zeros(X.shape[0], dtype=np.float32)
        check(self.lib, self.lib.lare_dev_sur_forward(_fp(X), X.shape[0], _fp(y)), "sur_forward")
        return y

    def grad(self, X, t):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(NPARAM, dtype=np.float32)
        loss = C.c_float()
        check(self.lib, self.lib.lare_dev_sur_grad(_fp(X), _fp(t), X.shape[0], _fp(g), C.byref(loss)), "sur_grad")
        return g, loss.value

    def enable(self, on=True, tol=1e-4):
        check(self.lib, self.lib.lare_dev_sur_enable(1 if on else 0, float(tol)), "sur_enable")

    def conduct(self):
        """One checked surrogate super-step of length dt on the device state:
        (residual, accepted)."""
        r = C.c_double()
        a = C.c_int32()
        check(self.lib, self.lib.lare_dev_sur_conduct(C.byref(r), C.byref(a)), "sur_conduct")
        return r.value, bool(a.value)

    def plain_residual(self):
        """The residual check applied to the plain RKL2 super-step from the
        current state (state unchanged): the floor of the check."""
        r = C.c_double()
        check(self.lib, self.lib.lare_dev_sur_plain_residual(C.byref(r)), "sur_plain_residual")
        return r.value

    def stats(self):
        a, rj = C.c_int32(), C.c_int32()
        r = C.c_double()
        check(self.lib, self.lib.lare_dev_sur_stats(C.byref(a), C.byref(rj), C.byref(r)), "sur_stats")
        return a.value, rj.value, r.value
