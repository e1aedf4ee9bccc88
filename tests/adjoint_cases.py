"""Seeded rollout states shared by the adjoint tests (CPU and GPU)."""
import numpy as np

GAMMA = 5.0 / 3.0


def smooth_field(rng, n1, n2, amp, kmax=3):
    x = (np.arange(n1 + 8) - 3.5) / n1
    y = (np.arange(n2 + 8) - 3.5) / n2
    X, Y = np.meshgrid(x, y)
    f = np.zeros_like(X)
    for _ in range(4):
        kx, ky = rng.integers(-kmax, kmax + 1, 2)
        f += rng.normal() * np.cos(2 * np.pi * (kx * X + ky * Y) + rng.uniform(0, 2 * np.pi))
    return amp * f


def rollout_state(seed, n1, n2, blast=False):
    """Dense (11, n2+8, n1+8) state: every field non-trivial, face B from a
    vector potential plus a uniform field, cell B = face averages."""
    rng = np.random.default_rng(seed)
    s = np.zeros((11, n2 + 8, n1 + 8))
    d = 1.0 + smooth_field(rng, n1, n2, 0.1)
    v = [smooth_field(rng, n1, n2, 0.3) for _ in range(3)]
    p = 1.0 + smooth_field(rng, n1, n2, 0.1)
    if blast:
        x = (np.arange(n1 + 8) - 3.5) / n1 - 0.5
        y = (np.arange(n2 + 8) - 3.5) / n2 - 0.5
        X, Y = np.meshgrid(x, y)
        p = np.where(X * X + Y * Y < 0.04, 10.0, 0.1)
    az = smooth_field(rng, n1, n2, 0.02)
    b1i = 0.5 + (np.roll(az, -1, axis=0) - az) * n2
    b2i = 0.2 - (np.roll(az, -1, axis=1) - az) * n1
    b3 = 0.3 + smooth_field(rng, n1, n2, 0.1)
    b1c = 0.5 * (b1i + np.roll(b1i, -1, axis=1))
    b2c = 0.5 * (b2i + np.roll(b2i, -1, axis=0))
    s[0] = d
    s[1], s[2], s[3] = d * v[0], d * v[1], d * v[2]
    s[4] = p / (GAMMA - 1) + 0.5 * d * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2) + \
        0.5 * (b1c ** 2 + b2c ** 2 + b3 ** 2)
    s[5], s[6], s[7] = b1c, b2c, b3
    s[8], s[9], s[10] = b1i, b2i, b3
    return s


def stable_dt(s, n1, n2, lx=1.0, ly=1.0, cfl=0.3):
    d = s[0]
    v = np.abs(s[1:4] / d).max()
    p = (GAMMA - 1) * (s[4] - 0.5 * (s[1] ** 2 + s[2] ** 2 + s[3] ** 2) / d -
                       0.5 * (s[5] ** 2 + s[6] ** 2 + s[7] ** 2))
    cf = np.sqrt((GAMMA * np.abs(p) + s[5] ** 2 + s[6] ** 2 + s[7] ** 2) / d).max()
    return cfl * min(lx / n1, ly / n2) / (v + cf)
