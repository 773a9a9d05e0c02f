"""CPU restatement of the learned wall model's training data, network and optimiser
-- TEST INFRASTRUCTURE ONLY.

The learned model has no reference code.  It sits at the reference's wall-model
callable boundary `model(gas, h, U, T_h, p_h, T_wall)` (`wmles/wall_model.py:3-9`,
called at `wmles/bc.py:161`) and is judged by the reference's own method: the
safeguarded fixed-count Newton solve for Spalding's law and Kader's temperature law
(`wmles/wall_model.py:19-61`, restated in `oracle.numerics.EquilibriumWallModel`).

* `uhash`, `dataset`: the sample generator (splitmix64 finaliser of (seed, index),
  features uniform over the trained box, targets log u+ and log T+ from the Newton
  oracle).  Pinned by: u+ satisfies Spalding's law (residual), T+ equals Kader's
  closed form, the hash's first outputs against the published splitmix64 sequence.
* `forward`, `loss_grad`: the 3 -> 64 -> 64 -> 2 network with the ISRU activation
  a / sqrt(1 + a^2) and the gradient of the
  mean squared error in float64.  Pinned by central finite differences of the loss
  and by torch.autograd (float64, CPU) on the same parameters.
* `adam`: Adam (Kingma & Ba 2015, Algorithm 1) with beta1 0.9, beta2 0.999, eps 1e-8.
  Pinned by its closed form on the first step (P - lr * sign(g) for |g| >> eps).
"""
import numpy as np

from .numerics import EquilibriumWallModel

H, NIN, NOUT = 64, 3, 2
# FP32 master layout: W1 [64][3], b1 [64], W2 [64][64], b2 [64], W3 [2][64], b3 [2]
O_W1 = 0
O_B1 = O_W1 + H * NIN
O_W2 = O_B1 + H
O_B2 = O_W2 + H * H
O_W3 = O_B2 + H
O_B3 = O_W3 + NOUT * H
NPARAM = O_B3 + NOUT

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    return (P[O_W1:O_B1].reshape(H, NIN), P[O_B1:O_W2], P[O_W2:O_B2].reshape(H, H),
            P[O_B2:O_W3], P[O_W3:O_B3].reshape(NOUT, H), P[O_B3:NPARAM])


def uhash(seed, i, s):
    """Uniform in [0, 1): splitmix64 finaliser of seed ^ (golden * (4 i + s + 1)) + golden."""
    with np.errstate(over="ignore"):
        g = np.uint64(0x9E3779B97F4A7C15)
        x = np.uint64(seed) ^ (g * (np.uint64(4) * np.asarray(i, dtype=np.uint64) + np.uint64(s + 1)))
        x = x + g
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def kader_tplus(Pr, yp):
    """Kader's T+ at y+ (wall_model.py:56-60)."""
    yp = np.maximum(yp, 1e-6)
    Gam = np.maximum(0.01 * (Pr * yp) ** 4 / (1.0 + 5.0 * Pr ** 3 * yp), 1e-20)
    beta = (3.85 * Pr ** (1 / 3) - 1.3) ** 2 + 2.12 * np.log(Pr)
    return Pr * yp * np.exp(-Gam) + (2.12 * np.log(1.0 + yp) + beta) * np.exp(-1.0 / Gam)


def uplus_newton(Reh, wm=None):
    """Spalding u+ = U/u_tau by the reference Newton (wall_model.py:19-40): U = 1,
    rho_w = 1, mu_w = 1/Re_h, h = 1."""
    wm = wm or EquilibriumWallModel()
    Reh = np.asarray(Reh, dtype=np.float64)
    return 1.0 / wm.utau(1.0, 1.0, Reh, 1.0)


def dataset(n, seed, lo, hi, x_center, x_scale, y_center, y_scale, Pr=0.72, wm=None):
    """Samples 0..n-1: (X float32 (n,3), Y float32 (n,2), raw float64 (n,5))."""
    i = np.arange(n, dtype=np.uint64)
    f = np.stack([lo[k] + (hi[k] - lo[k]) * uhash(seed, i, k) for k in range(3)], axis=1)
    Reh = np.exp(f[:, 0])
    up = uplus_newton(Reh, wm)
    Tp = kader_tplus(Pr, Reh / up)
    X = ((f - np.asarray(x_center, np.float64)) / np.asarray(x_scale, np.float64)).astype(np.float32)
    Y = np.stack([(np.log(up) - y_center[0]) / y_scale[0],
                  (np.log(Tp) - y_center[1]) / y_scale[1]], axis=1).astype(np.float32)
    raw = np.concatenate([f, up[:, None], Tp[:, None]], axis=1)
    return X, Y, raw


def act(a):
    """Hidden activation a / sqrt(1 + a^2)."""
    return a / np.sqrt(1.0 + a * a)


def act_prime(h):
    """Its derivative written in the activation h: (1 - h^2)^(3/2) = (1 + a^2)^(-3/2)."""
    return (1.0 - h * h) ** 1.5


def forward(P, X):
    """Network outputs (n, 2) in float64 and the hidden activations."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    X = np.asarray(X, dtype=np.float64)
    H1 = act(X @ W1.T + b1)
    H2 = act(H1 @ W2.T + b2)
    return H2 @ W3.T + b3, (X, H1, H2)


def loss_grad(P, X, Y):
    """Mean squared error over n x 2 entries and its gradient (float64, NPARAM)."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    out, (X, H1, H2) = forward(P, X)
    Y = np.asarray(Y, dtype=np.float64)
    n = X.shape[0]
    E = out - Y
    loss = float(np.sum(E * E) / (n * NOUT))
    dY = 2.0 * E / (n * NOUT)
    D2 = (dY @ W3) * act_prime(H2)
    D1 = (D2 @ W2) * act_prime(H1)
    g = np.zeros(NPARAM)
    g[O_W1:O_B1] = (D1.T @ X).ravel()
    g[O_B1:O_W2] = D1.sum(axis=0)
    g[O_W2:O_B2] = (D2.T @ H1).ravel()
    g[O_B2:O_W3] = D2.sum(axis=0)
    g[O_W3:O_B3] = (dY.T @ H2).ravel()
    g[O_B3:NPARAM] = dY.sum(axis=0)
    return loss, g


def adam(P, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1); returns (P, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return P - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed):
    """Xavier-uniform weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAM, dtype=np.float32)
    for off, fan_out, fan_in in ((O_W1, H, NIN), (O_W2, H, H), (O_W3, NOUT, H)):
        a = np.sqrt(6.0 / (fan_in + fan_out))
        P[off:off + fan_out * fan_in] = rng.uniform(-a, a, fan_out * fan_in).astype(np.float32)
    return P
