"""CPU FP64 oracle of the learned warm start.

TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker. The product
(lipika_amd/liblpk.so) never calls it.

The reference has no learned start (its solve always starts at t - 0.07 s,
src/channel/mod.rs:75-96), so this module restates the network this build
defines, in numpy float64:

  * the emission equation's light time on an analytic circular orbit
    (src/geometry/kepler.rs:32-77 with the Earth-fixed node; the linearised
    Sagnac rotation of src/channel/mod.rs:83-90), iterated to convergence;
  * the ten features and the target (tau - tau_sl) / 1e-6 s;
  * the network 10 -> 64 -> 64 -> 64 -> 1 (GELU, tanh form) with the BF16
    roundings of the matrix-core forward (inputs as a hi + lo pair, hidden
    weights and the activations of layers 1 and 2), or with none
    (quantize=False);
  * its mean-squared-error loss and the gradient by back-propagation, treating
    the roundings as the identity (straight-through).

Pins (tests/test_lws_oracle.py): the gradient against central finite
differences of the unrounded network; the light time against an independent
bisection of the emission equation with positions from the restated
reference propagator (orc_kepler_position); the feature parametrisation
against that propagator.
"""
from __future__ import annotations

import numpy as np

C_LIGHT = 299792458.0          # src/lib.rs:53
OMEGA_E = 7.2921151467e-5      # src/channel/mod.rs:27
GM_EARTH = 3.986005e14         # src/geometry/kepler.rs:12
R_EARTH = 6378137.0
A_MID, A_SCALE = 1.7e7, 1e7
TARGET_SCALE = 1e-6
GELU_K0 = np.float32(0.7978845608028654)
GELU_K1 = np.float32(0.044715)

N_IN, N_H = 10, 64
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3, OFF_W4, OFF_B4 = 0, 640, 704, 4800, 4864, 8960, 9024, 9088
N_PARAMS = 9089


def unpack(theta):
    t = np.asarray(theta, dtype=np.float64)
    return (t[OFF_W1:OFF_B1].reshape(N_H, N_IN), t[OFF_B1:OFF_W2], t[OFF_W2:OFF_B2].reshape(N_H, N_H),
            t[OFF_B2:OFF_W3], t[OFF_W3:OFF_B3].reshape(N_H, N_H), t[OFF_B3:OFF_W4], t[OFF_W4:OFF_B4], t[OFF_B4])


def bf16(x):
    """Round to the nearest BF16 (ties to even) through float32; returns float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def gelu(z):
    z = np.asarray(z, dtype=np.float64)
    return 0.5 * z * (1.0 + np.tanh(float(GELU_K0) * (z + float(GELU_K1) * z ** 3)))


def gelu_grad(z):
    z = np.asarray(z, dtype=np.float64)
    k0, k1 = float(GELU_K0), float(GELU_K1)
    th = np.tanh(k0 * (z + k1 * z ** 3))
    return 0.5 * (1.0 + th) + 0.5 * z * (1.0 - th * th) * k0 * (1.0 + 3.0 * k1 * z * z)


# ---------------------------------------------------------------- geometry

def orbit_position(a, inc, node, u):
    """Rz(node) Rx(inc) (a cos u, a sin u, 0): the operations of
    src/geometry/kepler.rs:41-48 with the Earth-fixed node."""
    a, inc, node, u = (np.asarray(v, dtype=np.float64) for v in (a, inc, node, u))
    xp, yp = a * np.cos(u), a * np.sin(u)
    y1, z1 = yp * np.cos(inc), yp * np.sin(inc)
    return np.stack([xp * np.cos(node) - y1 * np.sin(node), xp * np.sin(node) + y1 * np.cos(node), z1], axis=-1)


def _range_at(obs, a, inc, node, u, n, tau):
    p = orbit_position(a, inc, node + OMEGA_E * tau, u - n * tau)
    w = OMEGA_E * tau
    sx = p[..., 0] + w * p[..., 1]
    sy = p[..., 1] - w * p[..., 0]
    d = np.stack([sx, sy, p[..., 2]], axis=-1) - obs
    return np.sqrt((d * d).sum(axis=-1))


def light_time(obs, a, inc, node, u, iterations: int = 8):
    """(tau_sl, tau): straight-line light time and the converged root of
    tau = |R(w tau) P(tq - tau) - O| / c for the orbit in state (node, u) at tq."""
    obs = np.asarray(obs, dtype=np.float64)
    a = np.asarray(a, dtype=np.float64)
    n = np.sqrt(GM_EARTH / (a * a * a))
    p = orbit_position(a, inc, node, u)
    tau_sl = np.sqrt(((p - obs) ** 2).sum(axis=-1)) / C_LIGHT
    tau = tau_sl.copy()
    for _ in range(iterations):
        tau = _range_at(obs, a, inc, node, u, n, tau) / C_LIGHT
    return tau_sl, tau


def features(obs, a, inc, node, u):
    """The ten network inputs (float32), as the device forms them."""
    obs = np.asarray(obs, dtype=np.float64)
    cols = [obs[..., 0] / R_EARTH, obs[..., 1] / R_EARTH, obs[..., 2] / R_EARTH,
            (np.asarray(a, dtype=np.float64) - A_MID) / A_SCALE, np.cos(inc), np.sin(inc),
            np.cos(node), np.sin(node), np.cos(u), np.sin(u)]
    return np.stack([np.asarray(c, dtype=np.float64) for c in cols], axis=-1).astype(np.float32)


# ---------------------------------------------------------------- network

def _inputs(X, quantize):
    x = np.asarray(X, dtype=np.float32)
    if not quantize:
        return x.astype(np.float64)
    hi = bf16(x)
    lo = bf16(x - hi.astype(np.float32))
    return hi + lo


def forward(theta, X, quantize: bool = True):
    """Network output for each row of X [n, 10]; returns (y, cache)."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(theta)
    q = bf16 if quantize else (lambda v: np.asarray(v, dtype=np.float64))
    xe = _inputs(X, quantize)
    W1q, W2q, W3q = q(W1), q(W2), q(W3)
    z1 = xe @ W1q.T + b1
    h1 = gelu(z1)
    h1q = q(h1.astype(np.float32)) if quantize else h1
    z2 = h1q @ W2q.T + b2
    h2 = gelu(z2)
    h2q = q(h2.astype(np.float32)) if quantize else h2
    z3 = h2q @ W3q.T + b3
    h3 = gelu(z3)
    y = h3 @ w4 + b4
    return y, dict(xe=xe, z1=z1, h1q=h1q, z2=z2, h2q=h2q, z3=z3, h3=h3, W2q=W2q, W3q=W3q, w4=w4)


def loss_and_grad(theta, X, Y, quantize: bool = True):
    """Mean squared error over the rows and its gradient (float64 [N_PARAMS])."""
    y, c = forward(theta, X, quantize)
    n = len(y)
    d = y - np.asarray(Y, dtype=np.float64)
    loss = float((d * d).mean())
    dy = 2.0 * d / n
    g = np.zeros(N_PARAMS)
    g[OFF_W4:OFF_B4] = dy @ c["h3"]
    g[OFF_B4] = dy.sum()
    dz3 = dy[:, None] * c["w4"][None, :] * gelu_grad(c["z3"])
    g[OFF_W3:OFF_B3] = (dz3.T @ c["h2q"]).ravel()
    g[OFF_B3:OFF_W4] = dz3.sum(axis=0)
    dz2 = (dz3 @ c["W3q"]) * gelu_grad(c["z2"])
    g[OFF_W2:OFF_B2] = (dz2.T @ c["h1q"]).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(axis=0)
    dz1 = (dz2 @ c["W2q"]) * gelu_grad(c["z1"])
    g[OFF_W1:OFF_B1] = (dz1.T @ c["xe"]).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(axis=0)
    return loss, g
