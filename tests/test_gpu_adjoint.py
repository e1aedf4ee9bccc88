"""The adjoint kernels of the whole CTU+CT step chained over a checkpointed
fixed-dt rollout, against the tangent-linear model of the reference routines
(the oracle on dual numbers) by the dot-product identity
<ybar, J xdot> = <J^T ybar, xdot>, against central finite differences, and
through torch.autograd."""
import ctypes
import zlib
import importlib

import numpy as np
import pytest

from adjoint_cases import GAMMA, rollout_state, stable_dt
from helpers import load_pkg, oracle_lib, p

pytestmark = pytest.mark.gpu


def adj_mod():
    pkg = load_pkg()
    return importlib.import_module(pkg.__name__ + ".adjoint")


CASES = [
    # (name, n1, n2, flux, bc, nsteps, every, blast)
    ("roe_periodic", 32, 24, "roe", (4, 4, 4, 4), 6, 4, False),
    ("hlle_outflow", 24, 32, "hlle", (2, 2, 4, 4), 5, 2, False),
    ("roe_reflect_blast", 28, 28, "roe", (1, 1, 2, 1), 5, 3, True),
    ("hlle_blast_periodic", 24, 24, "hlle", (4, 4, 4, 4), 4, 4, True),
    # the smallest grid the adjoint takes (its ghost sources are interior)
    ("roe_smallest_mixed", 4, 5, "roe", (1, 2, 4, 4), 4, 2, False),
    # one segment: the forward pass stores every step, the reverse re-runs none
    ("hlle_one_segment", 20, 24, "hlle", (1, 4, 2, 4), 5, 5, True),
]
# three segments over two banks of slots: the reverse re-runs segment 0 into
# the bank that held the forward pass's last segment, so a second backward
# pass re-runs that segment too
DET_EXTRA = ("roe_reflect_blast_three_segments", 28, 28, "roe", (1, 1, 2, 1), 6, 2, True)


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_adjoint_rollout_dot_product_vs_tlm(case):
    import torch
    name, n1, n2, flux, bc, nsteps, every, blast = case
    print(f"[gpu] adjoint vs TLM {name} {n1}x{n2} N={nsteps}", flush=True)
    A = adj_mod()
    s0 = rollout_state(zlib.crc32(name.encode()) % 1000, n1, n2, blast=blast)
    dt = stable_dt(s0, n1, n2)
    desc = dict(nx1=n1, nx2=n2, x1min=0.0, x1max=1.0, x2min=0.0, x2max=1.0, gamma=GAMMA,
                bc=bc, flux=flux)
    h = A.Rollout(desc, dt, nsteps, every)
    rng = np.random.default_rng(1)
    xd = rng.normal(size=s0.shape)
    yb = rng.normal(size=s0.shape)
    O = oracle_lib()
    out, dout = np.zeros_like(s0), np.zeros_like(s0)
    O.ora_tlm_rollout(n1, n2, 0.0, 1.0, 0.0, 1.0, GAMMA, (ctypes.c_int * 4)(*bc),
                      0 if flux == "roe" else 1, dt, nsteps, p(s0), p(xd), p(out), p(dout))
    dev = torch.device("cuda:0")
    sN = h.forward(torch.from_numpy(s0).to(dev)).cpu().numpy()
    assert np.array_equal(sN, out), f"primal differs: {np.abs(sN - out).max()}"
    g0 = h.backward(torch.from_numpy(yb).to(dev)).cpu().numpy()
    lhs = yb.ravel() @ dout.ravel()
    rhs = g0.ravel() @ xd.ravel()
    print(f"[gpu] <ybar, J xdot> = {lhs:.15e}, <J^T ybar, xdot> = {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # ghost values of the input are overwritten by the first set_bvals
    ghost = np.ones(s0.shape[1:], bool)
    ghost[4:n2 + 4, 4:n1 + 4] = False
    assert np.all(g0[:, ghost] == 0.0) or np.all(g0[:8, ghost] == 0.0)
    h.close()


def _random_adjoint_cases(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n):
        n1, n2 = (int(x) for x in rng.choice([4, 5, 6, 9, 16, 24], size=2))
        bc = tuple(int(b) for b in rng.choice([1, 2, 4], size=4))
        nsteps = int(rng.integers(2, 7))
        every = int(rng.integers(1, nsteps + 1))
        out.append((f"random{seed}_{m}", n1, n2, str(rng.choice(["roe", "hlle"])), bc, nsteps,
                    every, bool(rng.integers(0, 2))))
    return out


@pytest.mark.parametrize("seed", [11, 12])
def test_adjoint_random_cases_dot_product_vs_tlm(seed):
    """A seeded sweep of rollouts: sizes from the smallest the adjoint takes,
    every combination of boundary kinds, either flux, 2 to 6 steps with any
    checkpoint spacing, smooth or blast states; each primal equals the
    oracle's and each adjoint satisfies the dot-product identity against the
    oracle's tangent-linear model to 1e-10."""
    for case in _random_adjoint_cases(seed, 6):
        test_adjoint_rollout_dot_product_vs_tlm(case)


def _fd_check(A, h, s0, target, dirs, eps_list=(1e-6,)):
    """max over directions of |FD - adjoint| / |adjoint| for each eps."""
    import torch
    dev = torch.device("cuda:0")
    n1, n2 = h.shape[2] - 8, h.shape[1] - 8
    tgt = torch.from_numpy(target).to(dev)
    inner = (slice(4, n2 + 4), slice(4, n1 + 4))

    def loss_of(state):
        return 0.5 * ((h.forward(state)[0][inner] - tgt) ** 2).sum()

    x = torch.from_numpy(s0).to(dev).requires_grad_(True)
    loss = 0.5 * ((A.rollout(x, h)[0][inner] - tgt) ** 2).sum()
    loss.backward()
    g = x.grad.detach()
    worst = {eps: 0.0 for eps in eps_list}
    for k, v in enumerate(dirs):
        v = torch.from_numpy(v).to(dev)
        ad = (g * v).sum().item()
        for eps in eps_list:
            fd = (loss_of(x.detach() + eps * v) - loss_of(x.detach() - eps * v)).item() / (2 * eps)
            rel = abs(fd - ad) / abs(ad)
            print(f"[gpu] direction {k} eps {eps:.0e}: fd {fd:.12e} adjoint {ad:.12e} "
                  f"rel {rel:.2e}", flush=True)
            worst[eps] = max(worst[eps], rel)
    return worst


def branch_digest(O, s, n, dt, nsteps, flux="roe", bc=(4, 4, 4, 4)):
    """Fingerprint of every branch the oracle's rollout from s takes (limiter
    and clamp selections, tracing signs, eigensystem degeneracies, Roe
    upwinding and HLLE fallback, corner-emf upwinding, |x|); the oracle's
    forward is the GPU forward bit for bit."""
    out, dout, z = np.zeros_like(s), np.zeros_like(s), np.zeros_like(s)
    dig, cnt = ctypes.c_ulonglong(), ctypes.c_ulonglong()
    O.ora_tlm_rollout_branches(n, n, 0.0, 1.0, 0.0, 1.0, GAMMA, (ctypes.c_int * 4)(*bc),
                               0 if flux == "roe" else 1, dt, nsteps, p(s), p(z), p(out),
                               p(dout), ctypes.byref(dig), ctypes.byref(cnt))
    return dig.value, cnt.value


def test_adjoint_gradient_vs_finite_differences_where_no_branch_flips():
    """A generic smooth state at 32^2, 16 steps (checkpoint every 4), loss =
    0.5 ||rho(T) - rho_target||^2, 32 seeded directions in all fields, central
    differences at eps = 1e-6, 1e-7, 1e-8. The step is piecewise smooth (the
    PLM limiter, clamps, upwinding, |lambda|, the HLLE fallback); for each
    direction and eps the oracle's branch fingerprints at x + eps v and
    x - eps v tell whether the difference quotient straddles a kink. Where
    nothing flips, FD must agree with the adjoint to 1e-6 relative (the
    path's bar for its gradients); directions that flip are reported."""
    import torch
    A = adj_mod()
    O = oracle_lib()
    n, nsteps, every = 32, 16, 4
    s0 = rollout_state(21, n, n)
    dt = stable_dt(s0, n, n)
    h = A.Rollout(dict(nx1=n, nx2=n, x1min=0.0, x1max=1.0, x2min=0.0, x2max=1.0, gamma=GAMMA,
                       bc=(4, 4, 4, 4), flux="roe"), dt, nsteps, every)
    target = rollout_state(22, n, n)[0][4:n + 4, 4:n + 4]
    dev = torch.device("cuda:0")
    tgt = torch.from_numpy(target).to(dev)
    inner = (slice(4, n + 4), slice(4, n + 4))

    def loss_of(state):
        return 0.5 * ((h.forward(state)[0][inner] - tgt) ** 2).sum().item()

    x = torch.from_numpy(s0).to(dev).requires_grad_(True)
    (0.5 * ((A.rollout(x, h)[0][inner] - tgt) ** 2).sum()).backward()
    grad = x.grad.detach()
    rng = np.random.default_rng(3)
    flat, flipped = [], {}
    for k in range(32):
        v_np = rng.normal(size=s0.shape)
        v = torch.from_numpy(v_np).to(dev)
        ad = (grad * v).sum().item()
        for eps in (1e-6, 1e-7, 1e-8):
            fd = (loss_of(x.detach() + eps * v) - loss_of(x.detach() - eps * v)) / (2 * eps)
            rel = abs(fd - ad) / abs(ad)
            flip = (branch_digest(O, s0 + eps * v_np, n, dt, nsteps) !=
                    branch_digest(O, s0 - eps * v_np, n, dt, nsteps))
            if flip:
                flipped.setdefault(eps, []).append(rel)
            else:
                flat.append((k, eps, rel))
    for eps in (1e-6, 1e-7, 1e-8):
        f = flipped.get(eps, [])
        ok = [r for _, e, r in flat if e == eps]
        print(f"[gpu] eps {eps:.0e}: {len(ok)} directions without a branch flip (max rel "
              f"{max(ok) if ok else float('nan'):.2e}), {len(f)} with flips (max rel "
              f"{max(f) if f else 0.0:.2e})", flush=True)
    h.close()
    assert len(flat) >= 32, "too few flip-free directions for the check to mean anything"
    worst = max(flat, key=lambda t: t[2])
    assert worst[2] <= 1e-6, f"direction {worst[0]} eps {worst[1]}: rel {worst[2]:.3e}"


def test_adjoint_gradient_vs_finite_differences_density():
    """The sensitivity problem at 64^2 (rho = 1 + 16 Fourier modes, P = 1,
    v = 0, uniform B): its velocity, pressure and field components sit on the
    limiter's kink (uniform data), so only density directions are
    differentiable there; checked along seeded random density directions at
    eps = 1e-6."""
    A = adj_mod()
    n, nsteps, every, dt = 64, 32, 8, 2e-4 * 8
    s0, _ = A.sensitivity_state(1, n)
    _, target = A.sensitivity_state(2, n)
    h = A.Rollout(A.sensitivity_desc(n), dt, nsteps, every)
    rng = np.random.default_rng(4)
    dirs = []
    for _ in range(4):
        v = np.zeros_like(s0)
        v[0] = rng.normal(size=s0.shape[1:])
        dirs.append(v)
    worst = _fd_check(A, h, s0, target, dirs)
    h.close()
    assert worst[1e-6] <= 1e-6


def test_superseded_rollout_backward_raises_on_the_device_handle():
    """Two rollouts on one handle before backward: the first graph's backward
    raises (the handle holds the second forward's checkpoints); the second
    graph's gradient equals a fresh handle's."""
    import torch
    A = adj_mod()
    n = 16
    s0 = rollout_state(5, n, n)
    s1 = rollout_state(6, n, n)
    dt = stable_dt(s0, n, n)
    desc = dict(nx1=n, nx2=n, x1min=0.0, x1max=1.0, x2min=0.0, x2max=1.0, gamma=GAMMA,
                bc=(4, 4, 4, 4), flux="roe")
    print("[gpu] stale rollout backward on a device handle", flush=True)
    dev = torch.device("cuda:0")
    h = A.Rollout(desc, dt, 4, 2)
    x0 = torch.from_numpy(s0).to(dev).requires_grad_(True)
    x1 = torch.from_numpy(s1).to(dev).requires_grad_(True)
    y0 = A.rollout(x0, h)
    y1 = A.rollout(x1, h)
    with pytest.raises(load_pkg().AthmiError, match="another forward pass"):
        y0[0].sum().backward()
    y1[0].sum().backward()
    h2 = A.Rollout(desc, dt, 4, 2)
    x1b = torch.from_numpy(s1).to(dev).requires_grad_(True)
    A.rollout(x1b, h2)[0].sum().backward()
    # the reverse scatters with atomic adds: two handles agree to roundoff of
    # the gradient's scale, not bit for bit
    err = (x1.grad - x1b.grad).abs().max().item()
    assert err <= 1e-12 * x1b.grad.abs().max().item(), err
    h.close()
    h2.close()


@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[5], DET_EXTRA], ids=lambda c: c[0])
def test_deterministic_reverse_is_bitwise_repeatable(case):
    """With the deterministic reverse two backward passes of one forward pass
    give bitwise equal gradients -- also where the first reads the last
    segment's slots as the forward pass filled them and the second re-runs
    that segment from its checkpoint; they agree with the default (atomic-add)
    reverse to roundoff and satisfy the dot-product identity against the TLM."""
    import torch
    name, n1, n2, flux, bc, nsteps, every, blast = case
    print(f"[gpu] deterministic reverse {name} {n1}x{n2} N={nsteps}", flush=True)
    A = adj_mod()
    s0 = rollout_state(zlib.crc32(name.encode()) % 1000, n1, n2, blast=blast)
    dt = stable_dt(s0, n1, n2)
    desc = dict(nx1=n1, nx2=n2, x1min=0.0, x1max=1.0, x2min=0.0, x2max=1.0, gamma=GAMMA,
                bc=bc, flux=flux)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(2)
    yb = rng.normal(size=s0.shape)
    xd = rng.normal(size=s0.shape)
    hd = A.Rollout(desc, dt, nsteps, every, deterministic=True)
    hd.forward(torch.from_numpy(s0).to(dev))
    g1 = hd.backward(torch.from_numpy(yb).to(dev)).cpu().numpy()
    g2 = hd.backward(torch.from_numpy(yb).to(dev)).cpu().numpy()
    assert np.array_equal(g1, g2), "deterministic backward passes differ"
    h = A.Rollout(desc, dt, nsteps, every)
    h.forward(torch.from_numpy(s0).to(dev))
    g0 = h.backward(torch.from_numpy(yb).to(dev)).cpu().numpy()
    assert np.abs(g1 - g0).max() <= 1e-12 * np.abs(g0).max()
    O = oracle_lib()
    out, dout = np.zeros_like(s0), np.zeros_like(s0)
    O.ora_tlm_rollout(n1, n2, 0.0, 1.0, 0.0, 1.0, GAMMA, (ctypes.c_int * 4)(*bc),
                      0 if flux == "roe" else 1, dt, nsteps, p(s0), p(xd), p(out), p(dout))
    lhs, rhs = yb.ravel() @ dout.ravel(), g1.ravel() @ xd.ravel()
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    hd.close()
    h.close()


def test_automatic_checkpoint_spacing_keeps_every_step_when_it_fits():
    """every=None: the library keeps every step (one segment, no re-run) when
    that fits in the free device memory, as it does for a small grid, and the
    gradient equals the checkpointed one (every 2) to roundoff."""
    import torch
    name, n1, n2, flux, bc, nsteps, every, blast = CASES[1]
    print(f"[gpu] automatic checkpoint spacing {name} {n1}x{n2} N={nsteps}", flush=True)
    A = adj_mod()
    s0 = rollout_state(zlib.crc32(name.encode()) % 1000, n1, n2, blast=blast)
    dt = stable_dt(s0, n1, n2)
    desc = dict(nx1=n1, nx2=n2, x1min=0.0, x1max=1.0, x2min=0.0, x2max=1.0, gamma=GAMMA,
                bc=bc, flux=flux)
    dev = torch.device("cuda:0")
    yb = torch.from_numpy(np.random.default_rng(3).normal(size=s0.shape)).to(dev)
    ha = A.Rollout(desc, dt, nsteps)
    assert ha.every == nsteps
    hk = A.Rollout(desc, dt, nsteps, 2)
    assert hk.every == 2
    sa = ha.forward(torch.from_numpy(s0).to(dev))
    sk = hk.forward(torch.from_numpy(s0).to(dev))
    assert torch.equal(sa, sk)
    ga, gk = ha.backward(yb), hk.backward(yb)
    assert (ga - gk).abs().max().item() <= 1e-12 * gk.abs().max().item()
    ha.close()
    hk.close()
