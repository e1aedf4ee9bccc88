"""The oracle restatement against the reference's own CPU sources compiled in
place (oracle/_ref), on fresh seeded inputs: routine by routine and whole
main-loop runs, value for value. Skipped where oracle/_ref is not built (it
is built from /root/reference by oracle/Makefile in the build container and
travels with the tree)."""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import INPUTS, REF_DIR, oracle_lib, oracle_setup, p, ref_available

pytestmark = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")
D, P = ctypes.c_double, ctypes.c_void_p


def ref_lib(flux="roe"):
    L = ctypes.CDLL(os.path.join(REF_DIR, f"libathref_{flux}.so"))
    L.ref_set_gamma.argtypes = [D]
    L.ref_flux_roe.argtypes = [P, P, P, P, D, P]
    L.ref_flux_hlle.argtypes = [P, P, P, P, D, P]
    L.ref_prim_to_cons.argtypes = [P, D, P]
    L.ref_cons_to_prim.argtypes = [P, D, P]
    return L


@pytest.mark.parametrize("gamma", [5.0 / 3.0, 1.4, 2.0])
def test_routines_match_reference_random(gamma):
    rng = np.random.default_rng(int(gamma * 1000))
    R, O = ref_lib(), oracle_lib()
    R.ref_set_gamma(gamma)
    for k in range(300):
        w = np.array([rng.uniform(0.05, 4), rng.normal(0, 2), rng.uniform(1e-3, 4),
                      rng.normal(0, 1.5), rng.normal(0, 1.5), rng.normal(0, 1.5)])
        if k % 17 == 0:
            w[4:6] = 0.0
        e1, r1, l1 = np.zeros(7), np.zeros(49), np.zeros(49)
        e2, r2, l2 = np.zeros(7), np.zeros(49), np.zeros(49)
        R.ref_esys_prim(p(w), p(e1), p(r1), p(l1))
        O.ora_esys_prim(gamma, p(w), p(e2), p(r2), p(l2))
        assert np.array_equal(e1, e2) and np.array_equal(r1, r2) and np.array_equal(l1, l2)
        Wl = np.array([rng.uniform(0.05, 4), *rng.normal(0, 2, 3), rng.uniform(1e-3, 4), *rng.normal(0, 1.5, 2)])
        Wr = np.array([rng.uniform(0.05, 4), *rng.normal(0, 2, 3), rng.uniform(1e-3, 4), *rng.normal(0, 1.5, 2)])
        bx = rng.normal(0, 1.5)
        Ul, Ur = np.zeros(7), np.zeros(7)
        R.ref_prim_to_cons(p(Wl), bx, p(Ul))
        R.ref_prim_to_cons(p(Wr), bx, p(Ur))
        for fr, fo in ((R.ref_flux_roe, O.ora_flux_roe), (R.ref_flux_hlle, O.ora_flux_hlle)):
            f1, f2 = np.zeros(7), np.zeros(7)
            fr(p(Ul), p(Ur), p(Wl), p(Wr), bx, p(f1))
            fo(gamma, p(Ul), p(Ur), p(Wl), p(Wr), bx, p(f2))
            assert np.array_equal(f1, f2), f"flux case {k}"


RUN = r"""
import ctypes, json, sys, numpy as np
L = ctypes.CDLL(sys.argv[1])
ov = json.loads(sys.argv[3])
arr = (ctypes.c_char_p * max(1, len(ov)))(*[o.encode() for o in ov])
L.ref_open(sys.argv[2].encode(), len(ov), arr)
L.ref_setup(int(sys.argv[4]))
iv = (ctypes.c_int * 9)(); dv = (ctypes.c_double * 6)(); L.ref_info(iv, dv)
sx, sy = iv[0] + 8, iv[1] + 8
L.ref_run(ctypes.c_double(float(sys.argv[5])), int(sys.argv[6]))
U = np.zeros((sy, sx, 8)); B = [np.zeros((sy, sx)) for _ in range(3)]
P = ctypes.c_void_p
L.ref_get(U.ctypes.data_as(P), *[b.ctypes.data_as(P) for b in B]); L.ref_info(iv, dv)
np.savez(sys.argv[7], gas=U, b1=B[0], b2=B[1], b3=B[2], clock=np.array([dv[4], dv[5], iv[6]]))
"""


def run_ref(tmp_path, flux, problem, overrides, tlim, nsteps):
    pid = {"blast": 0, "field_loop": 1, "cpaw2d": 2}[problem]
    out = str(tmp_path / "ref.npz")
    subprocess.run([sys.executable, "-c", RUN, os.path.join(REF_DIR, f"libathref_{flux}.so"),
                    os.path.join(INPUTS, f"athinput.{problem}"), json.dumps(overrides), str(pid),
                    str(tlim), str(nsteps), out], check=True, capture_output=True)
    return np.load(out)


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("flux", ["roe", "hlle"])
def test_main_loop_matches_reference_random_blast(tmp_path, seed, flux):
    rng = np.random.default_rng(seed)
    bcs = [int(rng.choice([1, 2, 4])) for _ in range(4)]
    if bcs[0] == 4 or bcs[1] == 4:
        bcs[0] = bcs[1] = 4
    if bcs[2] == 4 or bcs[3] == 4:
        bcs[2] = bcs[3] = 4
    ov = [f"grid/Nx1={int(rng.integers(20, 48))}", f"grid/Nx2={int(rng.integers(20, 48))}",
          f"grid/ibc_x1={bcs[0]}", f"grid/obc_x1={bcs[1]}", f"grid/ibc_x2={bcs[2]}",
          f"grid/obc_x2={bcs[3]}", f"problem/angle={rng.uniform(0, 180):.6f}",
          f"problem/prat={rng.uniform(2, 200):.6f}", f"problem/b0={rng.uniform(0, 2):.6f}",
          f"problem/radius={rng.uniform(0.1, 0.3):.6f}"]
    ref = run_ref(tmp_path, flux, "blast", ov, 0.2, 30)
    g, tlim, _ = oracle_setup("blast", ov, flux=flux)
    g.run(tlim, 30)
    gas, b1, b2, b3 = g.get()
    for name, x in (("gas", gas), ("b1", b1), ("b2", b2), ("b3", b3)):
        assert np.array_equal(x, ref[name]), f"{name}: {np.abs(x - ref[name]).max()}"
    assert list(g.clock()) == list(ref["clock"])


def test_cpaw2d_error_norm_matches_reference(tmp_path):
    """cpaw2d.c:222-302 prints the L1 errors with %e; the oracle's norm must
    agree to those digits after a run."""
    code = r"""
import ctypes, sys, os
L = ctypes.CDLL(sys.argv[1])
arr = (ctypes.c_char_p * 2)(b"grid/Nx1=32", b"grid/Nx2=32")
L.ref_open(sys.argv[2].encode(), 2, arr)
L.ref_setup(2)
L.ref_run(ctypes.c_double(1.0), 25)
L.ref_cpaw2d_errors(sys.argv[3].encode())
"""
    subprocess.run([sys.executable, "-c", code, os.path.join(REF_DIR, "libathref_roe.so"),
                    os.path.join(INPUTS, "athinput.cpaw2d"), str(tmp_path)], check=True,
                   capture_output=True)
    line = [l for l in open(tmp_path / "cpaw2d-errors.dat") if not l.startswith("#")][-1].split()
    ref = [float(x) for x in line[2:]]  # RMS, d, M1, M2, M3, E, B1c, B2c, B3c
    O = oracle_lib()
    g, tlim, _ = oracle_setup("cpaw2d", ["grid/Nx1=32", "grid/Nx2=32"])
    g.run(tlim, 25)
    out = np.zeros(9)
    O.ora_cpaw2d_l1(g.h, p(out))
    for a, b in zip(out, ref):
        assert float(f"{a:e}") == b
