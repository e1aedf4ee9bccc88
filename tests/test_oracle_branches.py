"""The oracle's branch fingerprint (ora_tlm_rollout_branches), which the GPU
finite-difference test uses to tell whether a central difference straddles a
kink of the piecewise-smooth step: it is reproducible, it separates
perturbations large enough to cross kinks, and where it reports no flip the
difference quotient agrees with the oracle's tangent-linear model (the exact
derivative along the branches taken) to 1e-6, while across flips it need
not."""
import ctypes

import numpy as np

from adjoint_cases import GAMMA, rollout_state, stable_dt
from helpers import oracle_lib, p

N, STEPS = 32, 16
INNER = (slice(4, N + 4), slice(4, N + 4))


def rollout(O, s, ds):
    out, dout = np.zeros_like(s), np.zeros_like(s)
    dig, cnt = ctypes.c_ulonglong(), ctypes.c_ulonglong()
    dt = stable_dt(rollout_state(21, N, N), N, N)
    O.ora_tlm_rollout_branches(N, N, 0.0, 1.0, 0.0, 1.0, GAMMA, (ctypes.c_int * 4)(4, 4, 4, 4),
                               0, dt, STEPS, p(s), p(ds), p(out), p(dout), ctypes.byref(dig),
                               ctypes.byref(cnt))
    return out, dout, dig.value, cnt.value


def test_fingerprint_is_reproducible_and_counts_every_decision():
    O = oracle_lib()
    s0 = rollout_state(21, N, N)
    z = np.zeros_like(s0)
    a, b = rollout(O, s0, z), rollout(O, s0, z)
    assert a[2] == b[2] and a[3] == b[3] and a[3] > 1_000_000
    # the primal of the logged run is the plain oracle rollout's
    out = np.zeros_like(s0)
    dout = np.zeros_like(s0)
    dt = stable_dt(s0, N, N)
    O.ora_tlm_rollout(N, N, 0.0, 1.0, 0.0, 1.0, GAMMA, (ctypes.c_int * 4)(4, 4, 4, 4), 0, dt,
                      STEPS, p(s0), p(z), p(out), p(dout))
    assert np.array_equal(out, a[0])


def test_flip_free_differences_agree_with_the_tangent_linear_model():
    O = oracle_lib()
    s0 = rollout_state(21, N, N)
    tgt = rollout_state(22, N, N)[0][INNER]
    rng = np.random.default_rng(5)

    def loss(out):
        return 0.5 * ((out[0][INNER] - tgt) ** 2).sum()

    smooth, kinked = [], []
    for _ in range(6):
        v = rng.normal(size=s0.shape)
        out, dout, _, _ = rollout(O, s0, v)
        tl = ((out[0][INNER] - tgt) * dout[0][INNER]).sum()
        for eps in (1e-6, 1e-8):
            op, _, dp, _ = rollout(O, s0 + eps * v, np.zeros_like(s0))
            om, _, dm, _ = rollout(O, s0 - eps * v, np.zeros_like(s0))
            rel = abs((loss(op) - loss(om)) / (2 * eps) - tl) / abs(tl)
            (kinked if dp != dm else smooth).append(rel)
    # eps 1e-6 crosses kinks in every direction here, 1e-8 in none
    assert len(smooth) >= 6 and len(kinked) >= 6
    assert max(smooth) <= 1e-6
    assert max(kinked) > 1e-6
