"""Write the golden fixtures of tests/golden/ from the reference's own CPU
sources compiled in place (oracle/_ref, built from /root/reference/src by
oracle/Makefile; this container only).

Fixtures (little-endian float64 .npz, plus JSON metadata):
  steps_<problem>_<flux>.npz   full padded state (Gas AoS, B1i, B2i, B3i) and
                               dt/time/nstep after k main-loop steps
  funcs.npz                    single-routine vectors: esys_prim, esys_roe,
                               flux_roe, flux_hlle, lr_states, conversions,
                               on seeded states including degenerate branches

Run: python tests/golden/make_golden.py   (needs oracle/_ref built)
"""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
INPUTS = os.path.join(ROOT, "athena-cuda_amd", "inputs")
D, I, P = ctypes.c_double, ctypes.c_int, ctypes.c_void_p

CASES = [  # problem, flux, overrides, checkpoints (main-loop steps)
    ("field_loop", "roe", [], [1, 10, 60]),
    ("field_loop", "hlle", [], [1, 10, 60]),
    ("blast", "roe", ["grid/Nx1=48", "grid/Nx2=72"], [1, 10, 60]),
    ("blast", "hlle", ["grid/Nx1=48", "grid/Nx2=72"], [1, 10, 60]),
    ("blast_outflow", "roe", ["grid/Nx1=40", "grid/Nx2=32", "grid/ibc_x1=2", "grid/obc_x1=2",
                              "grid/ibc_x2=2", "grid/obc_x2=2", "problem/radius=0.3"], [1, 20]),
    ("blast_reflect", "roe", ["grid/Nx1=40", "grid/Nx2=32", "grid/ibc_x1=1", "grid/obc_x1=1",
                              "grid/ibc_x2=1", "grid/obc_x2=1", "problem/radius=0.3"], [1, 20]),
    ("cpaw2d", "roe", ["grid/Nx1=32", "grid/Nx2=32"], [1, 10, 40]),
    ("cpaw2d", "hlle", ["grid/Nx1=32", "grid/Nx2=32"], [1, 10, 40]),
]
PROBLEM_ID = {"blast": 0, "field_loop": 1, "cpaw2d": 2}


def one_case(problem, flux, overrides, checkpoints):
    """Run in a fresh process: the reference keeps its boundary functions in
    file-scope statics, so each configuration needs its own library instance."""
    base = problem.split("_")[0] if problem.startswith("blast") else problem
    L = ctypes.CDLL(os.path.join(REF, f"libathref_{flux}.so"))
    ov = [o.encode() for o in overrides]
    arr = (ctypes.c_char_p * max(1, len(ov)))(*ov)
    L.ref_open(os.path.join(INPUTS, f"athinput.{base}").encode(), len(ov), arr)
    L.ref_setup(PROBLEM_ID[base])
    iv = (I * 9)()
    dv = (D * 6)()
    L.ref_info(iv, dv)
    sx, sy = iv[0] + 8, iv[1] + 8
    out = {}
    tlim = {"blast": 0.2, "field_loop": 2.0, "cpaw2d": 1.0}[base]

    def snap(tag):
        U = np.zeros((sy, sx, 8))
        B = [np.zeros((sy, sx)) for _ in range(3)]
        L.ref_get(U.ctypes.data_as(P), *[b.ctypes.data_as(P) for b in B])
        L.ref_info(iv, dv)
        out[f"{tag}_gas"] = U
        for n, b in zip(("b1i", "b2i", "b3i"), B):
            out[f"{tag}_{n}"] = b
        out[f"{tag}_clock"] = np.array([dv[4], dv[5], iv[6]], dtype=np.float64)

    snap("s0")
    done = 0
    for c in checkpoints:
        L.ref_run(D(tlim), c - done)
        done = c
        snap(f"s{c}")
    name = f"steps_{problem}_{flux}.npz"
    np.savez_compressed(os.path.join(HERE, name), **out)
    return {"file": name, "problem": base, "flux": flux, "overrides": overrides,
            "checkpoints": checkpoints, "shape": [sy, sx]}


def funcs():
    rng = np.random.default_rng(20261202)
    L = ctypes.CDLL(os.path.join(REF, "libathref_roe.so"))
    gamma = 5.0 / 3.0
    L.ref_set_gamma(D(gamma))
    out = {}
    n = 64
    # primitive states (d, vx, p, bx, by, bz) incl. degenerate branches
    prim = np.column_stack([rng.uniform(0.2, 3.0, n), rng.normal(0, 1, n), rng.uniform(0.05, 3, n),
                            rng.normal(0, 1, n), rng.normal(0, 1, n), rng.normal(0, 1, n)])
    prim[0, 4:6] = 0.0          # bt == 0
    prim[1, 3:6] = 0.0          # B == 0: cf2_cs2 == 0 hydro limit
    prim[2, 3] = 0.0            # bx == 0
    prim[3, 4] = 0.0
    prim[4, 3] = -0.0           # SIGN(-0) = +1
    ev = np.zeros((n, 7)); R = np.zeros((n, 49)); Lm = np.zeros((n, 49))
    for k in range(n):
        a = np.ascontiguousarray(prim[k])
        e = np.zeros(7); r = np.zeros(49); l = np.zeros(49)
        L.ref_esys_prim(a.ctypes.data_as(P), e.ctypes.data_as(P), r.ctypes.data_as(P), l.ctypes.data_as(P))
        ev[k], R[k], Lm[k] = e, r, l
    out.update(esys_prim_in=prim, esys_prim_ev=ev, esys_prim_R=R, esys_prim_L=Lm)
    # Roe-averaged states (d, v1, v2, v3, h, b1, b2, b3, x, y)
    roe = np.column_stack([rng.uniform(0.2, 3, n), rng.normal(0, 1, n), rng.normal(0, 1, n),
                           rng.normal(0, 1, n), rng.uniform(2, 8, n), rng.normal(0, 1, n),
                           rng.normal(0, 1, n), rng.normal(0, 1, n), rng.uniform(0, 0.2, n),
                           rng.uniform(1, 1.3, n)])
    roe[0, 6:8] = 0.0           # bt == 0
    roe[1, 5:8] = 0.0           # no field
    roe[2, 4] = 0.1             # twid_asq floor (h small)
    evr = np.zeros((n, 7)); Rr = np.zeros((n, 49)); Lr = np.zeros((n, 49)); ev_only = np.zeros((n, 7))
    for k in range(n):
        a = np.ascontiguousarray(roe[k])
        e = np.zeros(7); r = np.zeros(49); l = np.zeros(49)
        L.ref_esys_roe(a.ctypes.data_as(P), e.ctypes.data_as(P), r.ctypes.data_as(P), l.ctypes.data_as(P), 0)
        evr[k], Rr[k], Lr[k] = e, r, l
        e2 = np.zeros(7)
        L.ref_esys_roe(a.ctypes.data_as(P), e2.ctypes.data_as(P), r.ctypes.data_as(P), l.ctypes.data_as(P), 1)
        ev_only[k] = e2
    out.update(esys_roe_in=roe, esys_roe_ev=evr, esys_roe_R=Rr, esys_roe_L=Lr, esys_roe_evonly=ev_only)
    # Riemann problems: W (d, vx, vy, vz, p, by, bz) left/right and Bx
    m = 96
    def prims(scale):
        return np.column_stack([rng.uniform(0.1, 3, m), rng.normal(0, scale, m), rng.normal(0, 1, m),
                                rng.normal(0, 1, m), rng.uniform(0.01, 3, m), rng.normal(0, 1, m),
                                rng.normal(0, 1, m)])
    Wl, Wr = prims(1.0), prims(1.0)
    Wl[:16, 1] += 6.0; Wr[:16, 1] += 6.0          # supersonic to the right
    Wl[16:32, 1] -= 6.0; Wr[16:32, 1] -= 6.0      # supersonic to the left
    Wl[32:48, 1] -= 4.0; Wr[32:48, 1] += 4.0      # strong rarefaction: HLLE fallback
    Wl[48:56, 4] = 1e-3; Wr[48:56, 4] = 1e-3
    Wr[56:60] = Wl[56:60]                          # F(U,U) = F(U)
    bx = rng.normal(0, 1, m); bx[60:64] = 0.0
    Ul = np.zeros((m, 7)); Ur = np.zeros((m, 7)); Froe = np.zeros((m, 7)); Fhll = np.zeros((m, 7))
    for k in range(m):
        for W, U in ((Wl, Ul), (Wr, Ur)):
            w = np.ascontiguousarray(W[k]); u = np.zeros(7)
            L.ref_prim_to_cons(w.ctypes.data_as(P), D(bx[k]), u.ctypes.data_as(P))
            U[k] = u
        a = [np.ascontiguousarray(x[k]) for x in (Ul, Ur, Wl, Wr)]
        f = np.zeros(7)
        L.ref_flux_roe(*[x.ctypes.data_as(P) for x in a], D(bx[k]), f.ctypes.data_as(P))
        Froe[k] = f
        f2 = np.zeros(7)
        L.ref_flux_hlle(*[x.ctypes.data_as(P) for x in a], D(bx[k]), f2.ctypes.data_as(P))
        Fhll[k] = f2
    # conversion back (cons -> prim)
    Wback = np.zeros((m, 7))
    for k in range(m):
        u = np.ascontiguousarray(Ul[k]); w = np.zeros(7)
        L.ref_cons_to_prim(u.ctypes.data_as(P), D(bx[k]), w.ctypes.data_as(P))
        Wback[k] = w
    out.update(riemann_Wl=Wl, riemann_Wr=Wr, riemann_bx=bx, riemann_Ul=Ul, riemann_Ur=Ur,
               flux_roe=Froe, flux_hlle=Fhll, cons_to_prim=Wback)
    # lr_states on a slice of 40 cells with a shock and a smooth part
    nc = 40
    x = np.arange(nc)
    W = np.column_stack([1 + 0.5 * (x > 20) + 0.1 * np.sin(x / 3), 0.3 * np.cos(x / 5),
                         0.2 * np.sin(x / 4), 0.1 * np.ones(nc), 0.6 + 0.4 * (x > 20),
                         0.7 * np.cos(x / 6), 0.2 * np.ones(nc)])
    W[10:14, 1] = 0.0            # v = 0: ev[3] = 0 branch
    bxc = 0.8 + 0.05 * np.sin(x / 7)
    Wl_ = np.zeros((nc, 7)); Wr_ = np.zeros((nc, 7))
    W = np.ascontiguousarray(W)
    L.ref_lr_states(nc, W.ctypes.data_as(P), bxc.ctypes.data_as(P), D(0.01), D(0.4), 3, nc - 4,
                    Wl_.ctypes.data_as(P), Wr_.ctypes.data_as(P))
    out.update(lr_W=W, lr_bxc=bxc, lr_dt=np.array([0.01, 0.4]), lr_range=np.array([3, nc - 4]),
               lr_Wl=Wl_, lr_Wr=Wr_)
    np.savez_compressed(os.path.join(HERE, "funcs.npz"), **out)
    return {"file": "funcs.npz", "gamma": gamma}


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--case":
        k = int(sys.argv[2])
        print("RESULT " + json.dumps(one_case(*CASES[k])), flush=True)
        sys.exit(0)
    meta = {"source": "reference CPU sources compiled by oracle/Makefile (oracle/_ref)",
            "cases": [], "funcs": None}
    for k in range(len(CASES)):
        r = subprocess.run([sys.executable, __file__, "--case", str(k)], capture_output=True,
                           text=True, check=True)
        line = [x for x in r.stdout.splitlines() if x.startswith("RESULT ")][-1]
        meta["cases"].append(json.loads(line[len("RESULT "):]))
    r = subprocess.run([sys.executable, "-c", f"import sys; sys.path.insert(0, {HERE!r}); "
                        "import make_golden, json; print('RESULT ' + json.dumps(make_golden.funcs()), flush=True)"],
                       capture_output=True, text=True, check=True)
    line = [x for x in r.stdout.splitlines() if x.startswith("RESULT ")][-1]
    meta["funcs"] = json.loads(line[len("RESULT "):])
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    print("wrote", len(meta["cases"]), "step fixtures and funcs.npz")
