"""Checks that pin the oracle by mathematics rather than by the reference's
code: each would fail for a dropped term, a wrong sign or index, or a
transposed operand.

- esys_prim: L R = I and A r_k = lambda_k r_k with the closed-form Jacobian of
  the 1-D primitive ideal-MHD system (Stone et al. 2008, eq. A2).
- esys_roe + Roe averages: L R = I; the Roe property
  R diag(lambda) L (U_R - U_L) = F(U_R) - F(U_L) (Cargo & Gallice 1997) holds
  exactly for continuous transverse field, and to second order in the jump
  otherwise (a property of the reference's linearisation, not of the oracle:
  the oracle reproduces the reference routine bit for bit).
- fluxes: consistency F(U, U) = F(U); supersonic upwinding; the HLLE formula.
- lr_states: linear data with dt -> 0 gives W +- dW/2; constant data gives W.
- the step: a uniform state stays exactly uniform; periodic runs conserve
  mass, momentum and energy to roundoff; div B stays at roundoff.
- new_dt and set_bvals against brute-force numpy versions.
- cpaw2d: second-order L1 convergence (Gardiner & Stone 2005).
"""
import numpy as np
import pytest

from helpers import OracleGrid, load_pkg, oracle_lib, oracle_setup, p

G = 5.0 / 3.0


def prim_jacobian(d, vx, p_, bx, by, bz, gamma):
    A = np.zeros((7, 7))
    A[0, 0], A[0, 1] = vx, d
    A[1, 1], A[1, 4], A[1, 5], A[1, 6] = vx, 1 / d, by / d, bz / d
    A[2, 2], A[2, 5] = vx, -bx / d
    A[3, 3], A[3, 6] = vx, -bx / d
    A[4, 1], A[4, 4] = gamma * p_, vx
    A[5, 1], A[5, 2], A[5, 5] = by, -bx, vx
    A[6, 1], A[6, 3], A[6, 6] = bz, -bx, vx
    return A


def phys_flux(u, bx, gamma):
    d, mx, my, mz, e, by, bz = u
    vx, vy, vz = mx / d, my / d, mz / d
    pb = 0.5 * (bx * bx + by * by + bz * bz)
    pg = (gamma - 1) * (e - 0.5 * d * (vx * vx + vy * vy + vz * vz) - pb)
    pt = pg + pb
    return np.array([mx, mx * vx + pt - bx * bx, my * vx - bx * by, mz * vx - bx * bz,
                     (e + pt) * vx - bx * (bx * vx + by * vy + bz * vz),
                     by * vx - bx * vy, bz * vx - bx * vz])


def rand_prim(rng):
    return np.array([rng.uniform(0.3, 3), *rng.normal(0, 1, 3), rng.uniform(0.1, 2), *rng.normal(0, 1, 2)])


@pytest.mark.parametrize("seed", range(5))
def test_esys_prim_is_the_eigensystem_of_the_jacobian(seed):
    rng = np.random.default_rng(seed)
    L = oracle_lib()
    for _ in range(40):
        w = np.array([rng.uniform(0.3, 3), rng.normal(), rng.uniform(0.1, 2), *rng.normal(0, 1, 3)])
        ev, R, Lm = np.zeros(7), np.zeros(49), np.zeros(49)
        L.ora_esys_prim(G, p(w), p(ev), p(R), p(Lm))
        R, Lm = R.reshape(7, 7), Lm.reshape(7, 7)
        assert np.allclose(Lm @ R, np.eye(7), atol=1e-12)
        A = prim_jacobian(w[0], w[1], w[2], w[3], w[4], w[5], G)
        for k in range(7):
            assert np.allclose(A @ R[:, k], ev[k] * R[:, k], atol=1e-11 * (1 + abs(ev[k])))
        assert np.all(np.diff(ev) >= -1e-14)


def roe_inputs(Ul, Ur, Wl, Wr, bx):
    sl, sr = np.sqrt(Wl[0]), np.sqrt(Wr[0])
    i = 1 / (sl + sr)
    v = [(sl * Wl[k] + sr * Wr[k]) * i for k in (1, 2, 3)]
    b2 = (sr * Wl[5] + sl * Wr[5]) * i
    b3 = (sr * Wl[6] + sl * Wr[6]) * i
    x = 0.5 * ((Wl[5] - Wr[5]) ** 2 + (Wl[6] - Wr[6]) ** 2) / (sl + sr) ** 2
    y = 0.5 * (Wl[0] + Wr[0]) / (sl * sr)
    pbl = 0.5 * (bx ** 2 + Wl[5] ** 2 + Wl[6] ** 2)
    pbr = 0.5 * (bx ** 2 + Wr[5] ** 2 + Wr[6] ** 2)
    h = ((Ul[4] + Wl[4] + pbl) / sl + (Ur[4] + Wr[4] + pbr) / sr) * i
    return np.array([sl * sr, *v, h, bx, b2, b3, x, y])


def roe_residual(L, Wl, Wr, bx):
    Ul, Ur = np.zeros(7), np.zeros(7)
    L.ora_prim_to_cons(G, p(Wl), bx, p(Ul))
    L.ora_prim_to_cons(G, p(Wr), bx, p(Ur))
    a = roe_inputs(Ul, Ur, Wl, Wr, bx)
    ev, R, Lm = np.zeros(7), np.zeros(49), np.zeros(49)
    L.ora_esys_roe(G, p(a), p(ev), p(R), p(Lm), 0)
    R, Lm = R.reshape(7, 7), Lm.reshape(7, 7)
    assert np.allclose(Lm @ R, np.eye(7), atol=1e-11)
    dF = phys_flux(Ur, bx, G) - phys_flux(Ul, bx, G)
    return R @ np.diag(ev) @ Lm @ (Ur - Ul) - dF, dF


@pytest.mark.parametrize("seed", range(5))
def test_roe_matrix_has_the_roe_property(seed):
    """R diag(ev) L (U_R - U_L) = F_R - F_L exactly when the transverse field
    is continuous (any Bx, any hydrodynamic jump)."""
    rng = np.random.default_rng(100 + seed)
    L = oracle_lib()
    for _ in range(40):
        Wl, Wr = rand_prim(rng), rand_prim(rng)
        Wr[5:] = Wl[5:]
        res, dF = roe_residual(L, Wl, Wr, rng.normal())
        assert np.abs(res).max() <= 1e-12 * (1 + np.abs(dF).max())


@pytest.mark.parametrize("seed", range(3))
def test_roe_matrix_is_a_consistent_linearisation(seed):
    """With a jump in the transverse field the reference's Roe matrix (and so
    the oracle) does not satisfy the Roe property exactly; it is a consistent
    linearisation: the residual falls quadratically with the jump size."""
    rng = np.random.default_rng(200 + seed)
    L = oracle_lib()
    for _ in range(10):
        W0, dW = rand_prim(rng), rng.normal(0, 1, 7)
        bx = rng.normal()
        r = []
        for eps in (1e-2, 1e-3):
            res, _ = roe_residual(L, W0 - 0.5 * eps * dW, W0 + 0.5 * eps * dW, bx)
            r.append(np.abs(res).max())
        assert r[1] <= r[0] / 50 + 1e-13, r


def test_fluxes_consistency_upwinding_and_hlle_formula():
    rng = np.random.default_rng(7)
    L = oracle_lib()
    for _ in range(60):
        W = rand_prim(rng)
        bx = rng.normal()
        U = np.zeros(7)
        L.ora_prim_to_cons(G, p(W), bx, p(U))
        f = np.zeros(7)
        L.ora_flux_roe(G, p(U), p(U), p(W), p(W), bx, p(f))
        assert np.allclose(f, phys_flux(U, bx, G), rtol=1e-12, atol=1e-12)
        L.ora_flux_hlle(G, p(U), p(U), p(W), p(W), bx, p(f))
        assert np.allclose(f, phys_flux(U, bx, G), rtol=1e-12, atol=1e-12)
        # supersonic: the upwind state's physical flux
        Wl, Wr = W.copy(), rand_prim(rng)
        Wl[1] += 20.0
        Wr[1] += 20.0
        Ul, Ur = np.zeros(7), np.zeros(7)
        L.ora_prim_to_cons(G, p(Wl), bx, p(Ul))
        L.ora_prim_to_cons(G, p(Wr), bx, p(Ur))
        L.ora_flux_roe(G, p(Ul), p(Ur), p(Wl), p(Wr), bx, p(f))
        assert np.allclose(f, phys_flux(Ul, bx, G), rtol=1e-12, atol=1e-12)


def test_lr_states_linear_and_constant_data():
    L = oracle_lib()
    n = 24
    x = np.arange(n, dtype=float)
    base = np.array([1.0, 0.2, -0.1, 0.05, 0.8, 0.3, -0.2])
    slope = np.array([0.01, 0.002, 0.003, -0.001, 0.004, 0.002, 0.001])
    W = np.ascontiguousarray(base[None, :] + x[:, None] * slope[None, :])
    bxc = np.full(n, 0.7)
    Wl, Wr = np.zeros((n, 7)), np.zeros((n, 7))
    L.ora_lr_states(G, n, p(W), p(bxc), 0.0, 0.0, 3, n - 4, p(Wl), p(Wr))
    for i in range(3, n - 3):
        assert np.allclose(Wl[i + 1], W[i] + 0.5 * slope, atol=1e-13)
        assert np.allclose(Wr[i], W[i] - 0.5 * slope, atol=1e-13)
    Wc = np.ascontiguousarray(np.tile(base, (n, 1)))
    L.ora_lr_states(G, n, p(Wc), p(bxc), 0.01, 0.3, 3, n - 4, p(Wl), p(Wr))
    assert np.array_equal(Wl[3:n - 2], Wc[3:n - 2]) and np.array_equal(Wr[2:n - 3], Wc[2:n - 3])


def uniform_grid(flux, bcs=(4, 4, 4, 4)):
    d = dict(nx1=24, nx2=20, x1min=0.0, x1max=1.0, x2min=0.0, x2max=0.8, gamma=G, cour_no=0.4,
             bc=bcs, flux=flux)
    g = OracleGrid(d)
    gas, b1, b2, b3 = g.get()
    u = np.array([1.3, 0.2 * 1.3, -0.1 * 1.3, 0.05 * 1.3, 0.0, 0.6, 0.4, 0.2])
    u[4] = 0.9 / (G - 1) + 0.5 * 1.3 * (0.2 ** 2 + 0.1 ** 2 + 0.05 ** 2) + 0.5 * (0.36 + 0.16 + 0.04)
    gas[:] = u
    b1[:] = 0.6
    b2[:] = 0.4
    b3[:] = 0.2
    g.put(gas, b1, b2, b3)
    g.start()
    return g, gas.copy(), b1.copy(), b2.copy(), b3.copy()


@pytest.mark.parametrize("flux", ["roe", "hlle"])
def test_uniform_state_is_preserved_exactly(flux):
    g, gas, b1, b2, b3 = uniform_grid(flux)
    g.run(1e9, 10)
    a = g.get()
    for x, y in zip(a, (gas, b1, b2, b3)):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("problem,ov", [("blast", ["grid/Nx1=40", "grid/Nx2=60"]),
                                        ("field_loop", ["grid/Nx1=48", "grid/Nx2=24"]),
                                        ("cpaw2d", ["grid/Nx1=24", "grid/Nx2=24"])])
def test_periodic_runs_conserve_and_keep_divb_zero(problem, ov):
    g, tlim, _ = oracle_setup(problem, ov)
    (_, _, is_, ie, js, je, *_), (x10, x20, dx1, dx2, *_) = g.info()

    def totals():
        gas, b1, b2, b3 = g.get()
        sl = gas[js:je + 1, is_:ie + 1]
        return sl[..., :5].sum(axis=(0, 1)), b1, b2

    t0, _, _ = totals()
    for _ in range(25):
        g.step(tlim)
        t, b1, b2 = totals()
        scale = np.abs(g.get()[0][js:je + 1, is_:ie + 1, :5]).sum(axis=(0, 1))
        assert np.all(np.abs(t - t0) <= 1e-12 * scale + 1e-15), (t - t0) / scale
        div = (b1[js:je + 1, is_ + 1:ie + 2] - b1[js:je + 1, is_:ie + 1]) / dx1 + \
              (b2[js + 1:je + 2, is_:ie + 1] - b2[js:je + 1, is_:ie + 1]) / dx2
        bs = max(np.abs(b1).max(), np.abs(b2).max()) / min(dx1, dx2)
        assert np.abs(div).max() <= 1e-14 * bs


def test_new_dt_matches_brute_force():
    g, tlim, _ = oracle_setup("blast", ["grid/Nx1=30", "grid/Nx2=40"])
    g.run(tlim, 5)
    (nx1, nx2, is_, ie, js, je, nstep, *_), (_, _, dx1, dx2, dt_old, _) = g.info()
    gas = g.get()[0][js:je + 1, is_:ie + 1]
    d, m1, m2, m3, e = (gas[..., q] for q in range(5))
    di = 1.0 / d
    v1, v2, v3 = m1 * di, m2 * di, m3 * di
    pg = np.maximum((G - 1) * (e - 0.5 * d * (v1 * v1 + v2 * v2 + v3 * v3)), 1e-20)
    cs = np.sqrt(G * pg * di)
    mx = max(((np.abs(v1) + cs) / dx1).max(), ((np.abs(v2) + cs) / dx2).max())
    g.new_dt()
    dt = g.clock()[0]
    assert dt == min(2.0 * dt_old, 0.4 / mx)


@pytest.mark.parametrize("bcs", [(4, 4, 4, 4), (2, 2, 2, 2), (1, 1, 1, 1), (1, 2, 4, 4)])
def test_set_bvals_matches_index_map(bcs):
    rng = np.random.default_rng(sum(bcs))
    d = dict(nx1=9, nx2=7, x1min=0.0, x1max=1.0, x2min=0.0, x2max=1.0, gamma=G, cour_no=0.4,
             bc=bcs)
    g = OracleGrid(d)
    sy, sx = 15, 17
    gas, b1, b2, b3 = rng.normal(size=(sy, sx, 8)), *rng.normal(size=(3, sy, sx))
    g.put(gas, b1, b2, b3)
    g.set_bvals()
    out = g.get()
    is_, ie, js, je = 4, 12, 4, 10

    def mapx(i, f):
        if i < is_:
            k = is_ - i
            return {4: ie - (k - 1), 2: is_, 1: is_ + (k - 1)}[bcs[0]], bcs[0] == 1
        k = i - ie
        if f == "b1":
            return {4: is_ + (k - 1), 2: ie, 1: ie - (k - 2)}[bcs[1]], bcs[1] == 1
        return {4: is_ + (k - 1), 2: ie, 1: ie - (k - 1)}[bcs[1]], bcs[1] == 1

    def mapy(j, f):
        if j < js:
            k = js - j
            return {4: je - (k - 1), 2: js, 1: js + (k - 1)}[bcs[2]], bcs[2] == 1
        k = j - je
        if f == "b2":
            return {4: js + (k - 1), 2: je, 1: je - (k - 2)}[bcs[3]], bcs[3] == 1
        return {4: js + (k - 1), 2: je, 1: je - (k - 1)}[bcs[3]], bcs[3] == 1

    src = {"gas": gas, "b1": b1, "b2": b2, "b3": b3}
    res = dict(zip(("gas", "b1", "b2", "b3"), out))
    for f in src:
        iex = ie + 1 if f == "b1" else ie
        jey = je + 1 if f == "b2" else je
        for j in range(sy):
            for i in range(sx):
                gx, gy = i < is_ or i > iex, j < js or j > jey
                if gy:
                    sj, ny = mapy(j, f)
                    si, nx = mapx(i, f) if gx else (i, False)
                elif gx and js <= j <= jey:
                    sj, ny = j, False
                    si, nx = mapx(i, f)
                else:
                    assert np.array_equal(res[f][j, i], src[f][j, i])
                    continue
                want = src[f][sj, si].copy()
                if f == "gas":
                    if nx:
                        want[1] = -want[1]
                    if ny:
                        want[2] = -want[2]
                assert np.array_equal(res[f][j, i], want), (f, j, i)


@pytest.mark.parametrize("flux", ["roe", "hlle"])
def test_cpaw2d_converges_at_second_order(flux):
    O = oracle_lib()
    errs = []
    for n in (16, 32, 64):
        g, tlim, _ = oracle_setup("cpaw2d", [f"grid/Nx1={n}", f"grid/Nx2={n}"], flux=flux)
        g.run(tlim, -1)
        out = np.zeros(9)
        O.ora_cpaw2d_l1(g.h, p(out))
        errs.append(out[0])
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(orders > 1.7), orders
