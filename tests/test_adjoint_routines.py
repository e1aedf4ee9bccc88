"""The hand-written adjoints of single routines (host build of the adjoint
kernels' source) against the tangent-linear model of the reference routines
(the oracle instantiated on dual numbers): for random inputs x, tangents xdot
and cotangents ybar, <ybar, J xdot> (TLM) must equal <J^T ybar, xdot>
(adjoint). The TLM itself is pinned by central finite differences of the
oracle's double routines."""
import ctypes

import numpy as np
import pytest

from helpers import load_pkg, oracle_lib, p

D, P, I = ctypes.c_double, ctypes.c_void_p, ctypes.c_int
GAMMA = 5.0 / 3.0


def vjp_lib():
    L = load_pkg().lib()
    L.athmi_vjp_cons_to_prim.argtypes = [D, P, D, P, P, P]
    L.athmi_vjp_prim_to_cons.argtypes = [D, P, D, P, P, P]
    L.athmi_vjp_esys_prim.argtypes = [D, P, P, P, P, P]
    L.athmi_vjp_esys_roe.argtypes = [D, P, P, P, P, P, I]
    L.athmi_vjp_flux.argtypes = [D, I, P, P, P, P, D, P, P, P, P, P, P]
    L.athmi_vjp_flux.restype = I
    L.athmi_vjp_plm.argtypes = [D, P, P, P, D, D, P, P, P, P, P, P, P, P]
    return L


def close(a, b, tol=1e-11):
    scale = max(abs(a), abs(b), 1e-300)
    return abs(a - b) <= tol * scale or abs(a - b) < 1e-13


def rand_prim(rng):
    return np.array([rng.uniform(0.3, 3.0), *rng.normal(0, 1.0, 3), rng.uniform(0.05, 3.0),
                     *rng.normal(0, 1.0, 2)])


def prim_to_cons(w, bx, gamma=GAMMA):
    u = np.zeros(7)
    O = oracle_lib()
    O.ora_prim_to_cons(gamma, p(np.ascontiguousarray(w)), bx, p(u))
    return u


def test_conversions_dot_product():
    rng = np.random.default_rng(1)
    O, A = oracle_lib(), vjp_lib()
    for _ in range(200):
        w = rand_prim(rng)
        bx = rng.normal()
        u = prim_to_cons(w, bx)
        ud, bd, wb = rng.normal(size=7), rng.normal(), rng.normal(size=7)
        # cons_to_prim
        wv, wd = np.zeros(7), np.zeros(7)
        O.ora_tlm_cons_to_prim(GAMMA, p(u), p(ud), bx, bd, p(wv), p(wd))
        ub, bxb = np.zeros(7), D()
        A.athmi_vjp_cons_to_prim(GAMMA, p(u), bx, p(wb), p(ub), ctypes.byref(bxb))
        assert close(wb @ wd, ub @ ud + bxb.value * bd)
        # prim_to_cons
        wd2, ub2 = rng.normal(size=7), rng.normal(size=7)
        uv, udot = np.zeros(7), np.zeros(7)
        O.ora_tlm_prim_to_cons(GAMMA, p(w), p(wd2), bx, bd, p(uv), p(udot))
        wbb, bxb2 = np.zeros(7), D()
        A.athmi_vjp_prim_to_cons(GAMMA, p(w), bx, p(ub2), p(wbb), ctypes.byref(bxb2))
        assert close(ub2 @ udot, wbb @ wd2 + bxb2.value * bd)


@pytest.mark.parametrize("gamma", [5.0 / 3.0, 1.4])
def test_esys_prim_dot_product(gamma):
    rng = np.random.default_rng(2)
    O, A = oracle_lib(), vjp_lib()
    for k in range(300):
        x = np.array([rng.uniform(0.3, 3), rng.normal(), rng.uniform(0.05, 3), *rng.normal(0, 1, 3)])
        if k % 10 == 0:
            x[5] = 0.0  # bz = 0 (bet3 = 0)
        xd = rng.normal(size=6)
        ev, dev, R, dR, L, dL = (np.zeros(n) for n in (7, 7, 49, 49, 49, 49))
        O.ora_tlm_esys_prim(gamma, p(x), p(xd), p(ev), p(dev), p(R), p(dR), p(L), p(dL))
        evb, Rb, Lb = rng.normal(size=7), rng.normal(size=49), rng.normal(size=49)
        xb = np.zeros(6)
        A.athmi_vjp_esys_prim(gamma, p(x), p(Rb), p(Lb), p(evb), p(xb))
        assert close(evb @ dev + Rb @ dR + Lb @ dL, xb @ xd), k


def roe_inputs(rng, gamma):
    """(d, v1, v2, v3, h, b1, b2, b3, x, y) from two random states."""
    wl, wr = rand_prim(rng), rand_prim(rng)
    bx = rng.normal()
    ul, ur = prim_to_cons(wl, bx, gamma), prim_to_cons(wr, bx, gamma)
    sl, sr = np.sqrt(wl[0]), np.sqrt(wr[0])
    isd = 1 / (sl + sr)
    v = (sl * wl[1:4] + sr * wr[1:4]) * isd
    b2 = (sr * wl[5] + sl * wr[5]) * isd
    b3 = (sr * wl[6] + sl * wr[6]) * isd
    x = 0.5 * ((wl[5] - wr[5]) ** 2 + (wl[6] - wr[6]) ** 2) / (sl + sr) ** 2
    y = 0.5 * (wl[0] + wr[0]) / (sl * sr)
    pbl = 0.5 * (bx * bx + wl[5] ** 2 + wl[6] ** 2)
    pbr = 0.5 * (bx * bx + wr[5] ** 2 + wr[6] ** 2)
    h = ((ul[4] + wl[4] + pbl) / sl + (ur[4] + wr[4] + pbr) / sr) * isd
    return np.array([sl * sr, *v, h, bx, b2, b3, x, y])


@pytest.mark.parametrize("vectors", [1, 0])
def test_esys_roe_dot_product(vectors):
    rng = np.random.default_rng(3 + vectors)
    O, A = oracle_lib(), vjp_lib()
    for k in range(300):
        x = roe_inputs(rng, GAMMA)
        xd = rng.normal(size=10)
        ev, dev, R, dR, L, dL = (np.zeros(n) for n in (7, 7, 49, 49, 49, 49))
        O.ora_tlm_esys_roe(GAMMA, p(x), p(xd), p(ev), p(dev), p(R), p(dR), p(L), p(dL))
        evb = rng.normal(size=7)
        Rb, Lb = (rng.normal(size=49) * vectors for _ in range(2))
        xb = np.zeros(10)
        A.athmi_vjp_esys_roe(GAMMA, p(x), p(Rb), p(Lb), p(evb), p(xb), vectors)
        assert close(evb @ dev + Rb @ dR + Lb @ dL, xb @ xd), k


@pytest.mark.parametrize("kind", [0, 1])
def test_flux_dot_product(kind):
    rng = np.random.default_rng(10 + kind)
    O, A = oracle_lib(), vjp_lib()
    flags = set()
    for k in range(400):
        wl, wr = rand_prim(rng), rand_prim(rng)
        if k % 4 == 0:  # strong jumps: exercise the HLLE fallback of flux_roe
            wl[4] *= 0.01
            wr[1] += 3.0 * rng.choice([-1, 1])
        if k % 7 == 0:  # supersonic: the upwind branches
            wl[1] += 6.0 * rng.choice([-1, 1])
            wr[1] = wl[1]
        bx = rng.normal()
        ul, ur = prim_to_cons(wl, bx), prim_to_cons(wr, bx)
        d = [rng.normal(size=7) for _ in range(4)]
        bd = rng.normal()
        f, fd = np.zeros(7), np.zeros(7)
        flag = O.ora_tlm_flux(GAMMA, kind, p(ul), p(d[0]), p(ur), p(d[1]), p(wl), p(d[2]), p(wr),
                              p(d[3]), bx, bd, p(f), p(fd))
        fb = rng.normal(size=7)
        bars = [np.zeros(7) for _ in range(4)]
        bxb = D()
        flag2 = A.athmi_vjp_flux(GAMMA, kind, p(ul), p(ur), p(wl), p(wr), bx, p(fb),
                                 *[p(b) for b in bars], ctypes.byref(bxb))
        assert flag == flag2
        flags.add(flag)
        rhs = sum(b @ t for b, t in zip(bars, d)) + bxb.value * bd
        assert close(fb @ fd, rhs), (k, flag)
    if kind == 0:
        assert flags >= {0, 1} or flags >= {0, 2}, flags


@pytest.mark.parametrize("dtodx", [0.0, 0.3])
def test_plm_dot_product_and_values(dtodx):
    rng = np.random.default_rng(20)
    O, A = oracle_lib(), vjp_lib()
    n = 7
    for k in range(300):
        W = np.array([rand_prim(rng) for _ in range(n)])
        if k % 3 == 0:  # smooth data: the limiter's interior branches
            W = np.array([W[3] + 0.05 * j * rng.normal(size=7) for j in range(n)])
            W[:, 0] = np.abs(W[:, 0]) + 0.05
            W[:, 4] = np.abs(W[:, 4]) + 0.01
        bxc = rng.normal(size=n)
        Wd, bd = rng.normal(size=(n, 7)), rng.normal(size=n)
        Wl, dWl, Wr, dWr = (np.zeros((n, 7)) for _ in range(4))
        O.ora_tlm_lr_states(GAMMA, n, p(W), p(Wd), p(bxc), p(bd), dtodx, 4, 2, p(Wl), p(dWl),
                            p(Wr), p(dWr))
        lb, rb = rng.normal(size=7), rng.normal(size=7)
        out_l, out_r, wmb, w0b, wpb = (np.zeros(7) for _ in range(5))
        bxcb = D()
        A.athmi_vjp_plm(GAMMA, p(W[2]), p(W[3]), p(W[4]), bxc[3], dtodx, p(lb), p(rb), p(out_l),
                        p(out_r), p(wmb), p(w0b), p(wpb), ctypes.byref(bxcb))
        assert np.array_equal(out_l, Wl[4]) and np.array_equal(out_r, Wr[3]), k
        lhs = lb @ dWl[4] + rb @ dWr[3]
        rhs = wmb @ Wd[2] + w0b @ Wd[3] + wpb @ Wd[4] + bxcb.value * bd[3]
        assert close(lhs, rhs, 1e-10), k


def test_tlm_matches_finite_differences():
    """Pins the dual-number instantiation itself: the TLM of flux_roe and of
    lr_states agrees with central differences of the double oracle."""
    rng = np.random.default_rng(30)
    O = oracle_lib()
    eps = 1e-7
    for k in range(50):
        wl, wr = rand_prim(rng), rand_prim(rng)
        bx = rng.normal()
        wd = [rng.normal(size=7), rng.normal(size=7)]
        bd = rng.normal()

        def F(s):
            a, b = wl + s * wd[0], wr + s * wd[1]
            c = bx + s * bd
            ul, ur = prim_to_cons(a, c), prim_to_cons(b, c)
            f = np.zeros(7)
            O.ora_flux_roe(GAMMA, p(ul), p(ur), p(a), p(b), c, p(f))
            return f

        ul, ur = prim_to_cons(wl, bx), prim_to_cons(wr, bx)
        # tangent of U through prim_to_cons
        u1, ud1, u2, ud2 = (np.zeros(7) for _ in range(4))
        O.ora_tlm_prim_to_cons(GAMMA, p(wl), p(wd[0]), bx, bd, p(u1), p(ud1))
        O.ora_tlm_prim_to_cons(GAMMA, p(wr), p(wd[1]), bx, bd, p(u2), p(ud2))
        f, fd = np.zeros(7), np.zeros(7)
        O.ora_tlm_flux(GAMMA, 0, p(ul), p(ud1), p(ur), p(ud2), p(wl), p(wd[0]), p(wr), p(wd[1]),
                       bx, bd, p(f), p(fd))
        fdiff = (F(eps) - F(-eps)) / (2 * eps)
        assert np.allclose(fd, fdiff, rtol=1e-5, atol=1e-6 * (1 + np.abs(fd).max())), k


def test_tlm_rollout_matches_finite_differences():
    """Pins the rollout TLM (integrate_2d + set_bvals on dual numbers) against
    central differences of the same rollout in doubles."""
    from adjoint_cases import rollout_state, stable_dt
    O = oracle_lib()
    n1, n2, nsteps = 12, 10, 3
    s0 = rollout_state(5, n1, n2)
    dt = stable_dt(s0, n1, n2)
    bc = (ctypes.c_int * 4)(4, 4, 2, 2)
    rng = np.random.default_rng(6)
    xd = rng.normal(size=s0.shape)
    yb = rng.normal(size=s0.shape)

    def run(state, tangent):
        out, dout = np.zeros_like(state), np.zeros_like(state)
        O.ora_tlm_rollout(n1, n2, 0.0, 1.0, 0.0, 1.0, GAMMA, bc, 0, dt, nsteps,
                          p(np.ascontiguousarray(state)), p(np.ascontiguousarray(tangent)),
                          p(out), p(dout))
        return out, dout

    _, dout = run(s0, xd)
    eps = 1e-6
    fp, _ = run(s0 + eps * xd, np.zeros_like(s0))
    fm, _ = run(s0 - eps * xd, np.zeros_like(s0))
    fd = yb.ravel() @ ((fp - fm) / (2 * eps)).ravel()
    tl = yb.ravel() @ dout.ravel()
    assert abs(fd - tl) <= 1e-6 * abs(tl), (fd, tl)
