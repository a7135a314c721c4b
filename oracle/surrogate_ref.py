"""CPU restatement of the learned conduction surrogate (numpy).

TEST INFRASTRUCTURE ONLY: the checker for the device surrogate kernels
(lare3d_amd/csrc/surrogate.hpp).  Only tests/ load it.

The surrogate has no counterpart in the reference.  What this module pins:
  * bf16():        round-to-nearest-even FP32 -> BF16, the operand rounding
                   of the MFMA kernels (checked in the tests against bit
                   patterns fixed by the format);
  * features():    the 10 cell channels (temperature of conduct.f90:72-73
                   with zero ionisation energy, cell-centred B as the face
                   average of shock_viscosity, lagran.F90:551-553, and
                   centred temperature differences);
  * forward() / backward(): the MLP and the gradient of the mean squared
                   error (checked in the tests against central finite
                   differences in float64);
  * sample_cells(): the splitmix64 hash (Steele, Lea and Flood) the device
                   uses to draw training cells, and its activity-biased
                   choice among candidates.
"""
import math

import numpy as np

CH = 12            # channels stored per cell (10 used)
NCH = 10
KREAL = 27 * CH    # 324
K = 336            # padded to a multiple of 16
H = 64
OFF_W1 = 0
OFF_B1 = K * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1

_M64 = (1 << 64) - 1


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFFFFFF
    return ((u >> 16) << 16).astype(np.uint32).view(np.float32)


def unpack(w):
    w = np.asarray(w)
    return (w[OFF_W1:OFF_B1].reshape(K, H), w[OFF_B1:OFF_W2], w[OFF_W2:OFF_B2].reshape(H, H),
            w[OFF_B2:OFF_W3], w[OFF_W3:OFF_B3], w[OFF_B3])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3, [b3]])


def _erf(x):
    return np.vectorize(math.erf)(x)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def forward(w, X, emulate_bf16=True, cache=False):
    """y = W3 . GELU(W2^T GELU(W1^T x + b1) + b2) + b3 for rows X (n x K).
    With emulate_bf16, GEMM operands are rounded to BF16 as on the device
    (the output layer stays FP32 there too)."""
    W1, b1, W2, b2, W3, b3 = unpack(w)
    dt = np.float32 if emulate_bf16 else np.float64
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    X = np.asarray(X, dtype=dt)
    Z1 = (r(X).astype(np.float64) @ r(W1).astype(np.float64)).astype(dt) + b1.astype(dt)
    H1 = gelu(Z1).astype(dt)
    Z2 = (r(H1).astype(np.float64) @ r(W2).astype(np.float64)).astype(dt) + b2.astype(dt)
    H2 = gelu(Z2).astype(dt)
    y = (H2.astype(np.float64) @ W3.astype(np.float64) + float(b3)).astype(dt)
    if cache:
        return y, (X, Z1, H1, Z2, H2)
    return y


def backward(w, X, t, emulate_bf16=True):
    """Gradient of mean((y - t)^2) with respect to the parameter vector, and the loss."""
    W1, b1, W2, b2, W3, b3 = unpack(w)
    dt = np.float32 if emulate_bf16 else np.float64
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    y, (X, Z1, H1, Z2, H2) = forward(w, X, emulate_bf16, cache=True)
    n = X.shape[0]
    e = y.astype(np.float64) - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(e * e))
    dy = (2.0 * e / n).astype(dt)
    gW3 = H2.astype(np.float64).T @ dy.astype(np.float64)
    gb3 = float(np.sum(dy, dtype=np.float64))
    dZ2 = (np.outer(dy, W3).astype(dt) * gelu_grad(Z2).astype(dt)).astype(dt)
    gb2 = dZ2.astype(np.float64).sum(axis=0)
    gW2 = r(H1).astype(np.float64).T @ r(dZ2).astype(np.float64)
    dH1 = (r(dZ2).astype(np.float64) @ r(W2).astype(np.float64).T).astype(dt)
    dZ1 = (dH1 * gelu_grad(Z1).astype(dt)).astype(dt)
    gb1 = dZ1.astype(np.float64).sum(axis=0)
    gW1 = r(X).astype(np.float64).T @ r(dZ1).astype(np.float64)
    return pack(gW1, gb1, gW2, gb2, gW3, gb3), loss


def init_params(seed=0):
    """Seeded initial parameters: W1 ~ N(0, 1/324) on the rows of the ten used
    channels (zero elsewhere), W2 ~ N(0, 1/64), W3 ~ N(0, 0.01/64), zero biases."""
    rng = np.random.default_rng(seed)
    W1 = np.zeros((K, H))
    used = np.array([(kk < KREAL) and (kk % CH < NCH) for kk in range(K)])
    W1[used] = rng.standard_normal((int(used.sum()), H)) / math.sqrt(27 * NCH)
    W2 = rng.standard_normal((H, H)) / math.sqrt(H)
    W3 = 0.1 * rng.standard_normal(H) / math.sqrt(H)
    return pack(W1, np.zeros(H), W2, np.zeros(H), W3, 0.0).astype(np.float32)


def mix64(x):
    """splitmix64 finaliser on Python ints."""
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def _cell_of(h, n):
    return (1 + (h & 0xFFFFF) % n[0], 1 + ((h >> 20) & 0xFFFFF) % n[1], 1 + ((h >> 40) & 0xFFFFF) % n[2])


def sample_cells(seed, rows, n, temperature=None):
    """Interior cells (i, j, k), 1-based, that the device draws for rows
    0..rows-1: even rows uniform; odd rows the candidate of 8 with the
    largest |6 T - sum of the six face neighbours| (FP32 arithmetic on the
    FP32 temperature channel; first maximum wins).  temperature(i, j, k)
    returns that channel; without it only even rows are valid."""
    out = []
    for r in range(rows):
        h = mix64((seed * 0x100000001B3 + r) & _M64)
        best = _cell_of(h, n)
        if r & 1:
            if temperature is None:
                best = None
            else:
                score = np.float32(-1.0)
                for t in range(8):
                    i, j, k = _cell_of(mix64((h + t + 1) & _M64), n)
                    lap = np.float32(6.0) * temperature(i, j, k)
                    for (a, b, c) in ((i + 1, j, k), (i - 1, j, k), (i, j + 1, k), (i, j - 1, k), (i, j, k + 1),
                                      (i, j, k - 1)):
                        lap = np.float32(lap - temperature(a, b, c))
                    sc = np.float32(abs(lap))
                    if sc > score:
                        score, best = sc, (i, j, k)
        out.append(best)
    return out


def features(grid, fields, dt, gamma, xi_n, cells):
    """Raw input rows (len(cells) x K) for the given interior cells.

    fields: a lare3d_amd.fields.Fields (Fortran bounds); grid: lare3d_amd.grid.Grid."""
    tfac = (gamma - 1.0) / (2.0 - xi_n)
    e = fields.energy
    lo_c = -1
    def E(i, j, k):
        return e[i - lo_c, j - lo_c, k - lo_c]
    def B(i, j, k):
        bx = 0.5 * (fields.bx[i + 2, j + 1, k + 1] + fields.bx[i + 1, j + 1, k + 1])
        by = 0.5 * (fields.by[i + 1, j + 2, k + 1] + fields.by[i + 1, j + 1, k + 1])
        bz = 0.5 * (fields.bz[i + 1, j + 1, k + 2] + fields.bz[i + 1, j + 1, k + 1])
        return bx, by, bz
    dxc = [grid.dxc[a] for a in range(3)]   # index [i + 1] for i = -1..n+1
    def chan(i, j, k):
        T = tfac * E(i, j, k)
        bx, by, bz = B(i, j, k)
        bm = math.sqrt(bx * bx + by * by + bz * bz)
        inv = 1.0 / max(bm, 1e-30)
        gx = tfac * (E(i + 1, j, k) - E(i - 1, j, k)) / (dxc[0][i + 1] + dxc[0][i])
        gy = tfac * (E(i, j + 1, k) - E(i, j - 1, k)) / (dxc[1][j + 1] + dxc[1][j])
        gz = tfac * (E(i, j, k + 1) - E(i, j, k - 1)) / (dxc[2][k + 1] + dxc[2][k])
        rho = fields.rho[i + 1, j + 1, k + 1]
        return np.array([T, bx * inv, by * inv, bz * inv, rho, math.log(dt), bm, gx, gy, gz, 0.0, 0.0],
                        dtype=np.float64).astype(np.float32)
    X = np.zeros((len(cells), K), dtype=np.float32)
    for r, (i, j, k) in enumerate(cells):
        for q in range(27):
            dx, dy, dz = q % 3 - 1, (q // 3) % 3 - 1, q // 9 - 1
            X[r, q * CH:(q + 1) * CH] = chan(i + dx, j + dy, k + dz)
    return X
