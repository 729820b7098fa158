"""FP64 numpy oracle of the learned Newton-guess network and its data.

Test infrastructure only: tests/ use it to check the HIP forward, backward
and data-generation kernels; nothing in the product imports it.

The network has no counterpart in the reference.  What it wraps does:
labels are star pressures of ExactRiemannSolver::ComputeMiddleState
(include/fub/equations/perfect_gas/ExactRiemannSolver.hpp:34-97) seeded with
the reference guess 0.5 (pL + pR) (ExactRiemannSolver.hpp:91), computed here
by the C oracle (oracle_exact_middle_state in fvs_oracle.cpp).

Parity pinning: `loss_and_grad` is pinned by central finite differences of
`loss` (tests/test_mlp_oracle.py); `forward` by a hand-evaluated one-unit
network; the sample generator by the closed-form ranges and the SplitMix64
reference sequence.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HIDDEN = 6, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_HIDDEN * 8
OFF_W2 = OFF_B1 + N_HIDDEN
OFF_B2 = OFF_W2 + N_HIDDEN * N_HIDDEN
OFF_W3 = OFF_B2 + N_HIDDEN
OFF_B3 = OFF_W3 + N_HIDDEN
NUM_PARAMS = OFF_B3 + 1
LOG_SCALE = 1.0 / np.log(1000.0)
MASK64 = (1 << 64) - 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(N_HIDDEN, 8)[:, :N_IN], p[OFF_B1:OFF_W2],
            p[OFF_W2:OFF_B2].reshape(N_HIDDEN, N_HIDDEN), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p, x):
    """x: (6, n) -> (y (n), h1 (64, n), h2 (64, n)) in FP64."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.maximum(W1 @ x + b1[:, None], 0.0)
    h2 = np.maximum(W2 @ h1 + b2[:, None], 0.0)
    y = W3 @ h2 + b3
    return y, h1, h2


def _pointwise(e, huber):
    """Per-sample loss: e^2, or with huber h > 0 the Huber form 2 h |e| - h^2
    beyond |e| > h (continuous, derivative 2 clamp(e, -h, h))."""
    if huber > 0:
        ae = np.abs(e)
        return np.where(ae <= huber, e * e, 2.0 * huber * ae - huber * huber)
    return e * e


def loss(p, x, t, flag, huber=0.0):
    """Mean per-sample loss over the samples with flag == 0."""
    y, _, _ = forward(p, x)
    ok = np.asarray(flag) == 0
    if not ok.any():
        return 0.0
    e = (y - np.asarray(t, dtype=np.float64))[ok]
    return float(np.sum(_pointwise(e, huber)) / ok.sum())


def loss_and_grad(p, x, t, flag, huber=0.0):
    """Loss and its gradient with respect to the flat parameter vector
    (padded W1 columns 6 and 7 get zero gradient)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    y, h1, h2 = forward(p, x)
    ok = (np.asarray(flag) == 0).astype(np.float64)
    cnt = max(ok.sum(), 1.0)
    e = (y - np.asarray(t, dtype=np.float64)) * ok
    ec = np.clip(e, -huber, huber) if huber > 0 else e
    dy = 2.0 * ec / cnt
    dz2 = (W3[:, None] * dy[None, :]) * (h2 > 0.0)
    dz1 = (W2.T @ dz2) * (h1 > 0.0)
    g = np.zeros(NUM_PARAMS)
    gW1 = np.zeros((N_HIDDEN, 8))
    gW1[:, :N_IN] = dz1 @ x.T
    g[OFF_W1:OFF_B1] = gW1.ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(axis=1)
    g[OFF_W2:OFF_B2] = (dz2 @ h1.T).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(axis=1)
    g[OFF_W3:OFF_B3] = h2 @ dy
    g[OFF_B3] = dy.sum()
    return float(np.sum(_pointwise(e, huber) * ok) / cnt), g


def features(prim):
    """prim (6, n) = (rhoL, uL, pL, rhoR, uR, pR) -> FP32 features as the
    device computes them (logf of the float-rounded value)."""
    prim = np.asarray(prim, dtype=np.float64)
    f = np.empty((N_IN, prim.shape[1]), dtype=np.float32)
    s = np.float32(LOG_SCALE)
    for side in range(2):
        rho, u, p = prim[3 * side], prim[3 * side + 1], prim[3 * side + 2]
        f[3 * side] = np.log(rho.astype(np.float32)) * s
        f[3 * side + 1] = ((rho * u) / rho).astype(np.float32) * np.float32(0.2)
        f[3 * side + 2] = np.log(p.astype(np.float32)) * s
    return f


def splitmix(x: int) -> int:
    """SplitMix64 finaliser (Steele, Lea, Flood: Fast splittable PRNGs, 2014)."""
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK64
    return x ^ (x >> 31)


def uniform(seed: int, i: int, k: int) -> float:
    return (splitmix(seed ^ splitmix(i * 8 + k)) >> 11) * (1.0 / 9007199254740992.0)


def sample_prim(seed: int, idx) -> np.ndarray:
    """The sample generator's primitive pairs for sample indices idx:
    rho, p log-uniform in [1e-3, 1e3], u uniform in [-5, 5]."""
    lo, hi = np.log(1e-3), np.log(1e3)
    out = np.empty((6, len(idx)))
    for c, i in enumerate(idx):
        u = [uniform(seed, int(i), k) for k in range(6)]
        out[:, c] = (np.exp(lo + (hi - lo) * u[0]), -5.0 + 10.0 * u[1],
                     np.exp(lo + (hi - lo) * u[2]), np.exp(lo + (hi - lo) * u[3]),
                     -5.0 + 10.0 * u[4], np.exp(lo + (hi - lo) * u[5]))
    return out


def complete_pairs(prim, gamma=1.4):
    """(6, n) primitive pairs -> 2D complete states (6, n) left and right:
    rho, rho u, 0, E, p, a."""
    prim = np.asarray(prim, dtype=np.float64)
    out = []
    for side in range(2):
        rho, u, p = prim[3 * side], prim[3 * side + 1], prim[3 * side + 2]
        m = rho * u
        q = np.empty((6, prim.shape[1]))
        q[0], q[1], q[2] = rho, m, 0.0
        q[3] = 0.5 * (m * m + 0.0) / rho + p * (1.0 / (gamma - 1.0))
        q[4] = p
        q[5] = np.sqrt(gamma * p / rho)
        out.append(q)
    return out[0], out[1]


def vacuum(prim, gamma=1.4):
    """Pressure-positivity condition of the exact solver: a vacuum forms when
    2 aL / (g - 1) + 2 aR / (g - 1) <= uR - uL (Toro, Riemann Solvers, 4.40)."""
    prim = np.asarray(prim, dtype=np.float64)
    gm1 = gamma - 1.0
    aL = np.sqrt(gamma * prim[2] / prim[0])
    aR = np.sqrt(gamma * prim[5] / prim[3])
    return 2.0 * aL / gm1 + 2.0 * aR / gm1 <= prim[4] - prim[1]
