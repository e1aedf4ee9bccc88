"""Test helpers: load the package (its directory name has a hyphen), and
ctypes wrappers of the checkers -- the oracle restatement (oracle/liboracle.so)
and the reference's own CPU sources compiled in place (oracle/_ref/*.so)."""
import ctypes
import importlib.util
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "athena-cuda_amd")
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
INPUTS = os.path.join(PKG_DIR, "inputs")

_D, _I, _P = ctypes.c_double, ctypes.c_int, ctypes.c_void_p


def load_pkg():
    name = "athena_cuda_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def ensure_oracle():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       stdout=subprocess.DEVNULL)
    return ORACLE_SO


def ref_available(flux="roe"):
    return os.path.exists(os.path.join(REF_DIR, f"libathref_{flux}.so"))


_ora = None


def oracle_lib():
    global _ora
    if _ora is None:
        L = ctypes.CDLL(ensure_oracle())
        L.ora_grid_new.restype = _P
        L.ora_grid_new.argtypes = [_I, _I, _D, _D, _D, _D, _D, _D, _I, _I, _I, _I, _I]
        for n in ("ora_grid_free", "ora_start", "ora_integrate", "ora_new_dt", "ora_set_bvals"):
            getattr(L, n).argtypes = [_P]
        L.ora_step.argtypes = [_P, _D]
        L.ora_run.argtypes = [_P, _D, _I]
        L.ora_info.argtypes = [_P, _P, _P]
        L.ora_get.argtypes = [_P, _P, _P, _P, _P]
        L.ora_put.argtypes = [_P, _P, _P, _P, _P]
        L.ora_set_time.argtypes = [_P, _D, _D, _I]
        L.ora_pgen_blast.argtypes = [_P, _D, _D, _D, _D, _D]
        for name in ("ora_tlm_cons_to_prim", "ora_tlm_prim_to_cons"):
            getattr(L, name).argtypes = [_D, _P, _P, _D, _D, _P, _P]
        L.ora_tlm_esys_prim.argtypes = [_D] + [_P] * 8
        L.ora_tlm_esys_roe.argtypes = [_D] + [_P] * 8
        L.ora_tlm_flux.argtypes = [_D, _I] + [_P] * 8 + [_D, _D, _P, _P]
        L.ora_tlm_flux.restype = _I
        L.ora_tlm_lr_states.argtypes = [_D, _I, _P, _P, _P, _P, _D, _I, _I, _P, _P, _P, _P]
        L.ora_tlm_rollout.argtypes = [_I, _I, _D, _D, _D, _D, _D, _P, _I, _D, _I, _P, _P, _P, _P]
        L.ora_tlm_rollout_branches.argtypes = [_I, _I, _D, _D, _D, _D, _D, _P, _I, _D, _I, _P, _P,
                                               _P, _P, _P, _P]
        L.ora_pgen_blast_at.argtypes = [_P, _D, _D, _D, _D, _D, _D, _D]
        L.ora_pgen_field_loop.argtypes = [_P, _D, _D, _D, _I]
        L.ora_pgen_cpaw2d.argtypes = [_P, _D, _D, _D, _D, _D]
        L.ora_pgen_cpaw2d.restype = _D
        L.ora_cpaw2d_l1.argtypes = [_P, _P]
        L.ora_max_inverse_dt.argtypes = [_P]
        L.ora_max_inverse_dt.restype = _D
        L.ora_esys_prim.argtypes = [_D, _P, _P, _P, _P]
        L.ora_esys_roe.argtypes = [_D, _P, _P, _P, _P, _I]
        L.ora_flux_roe.argtypes = [_D, _P, _P, _P, _P, _D, _P]
        L.ora_flux_roe.restype = _I
        L.ora_flux_hlle.argtypes = [_D, _P, _P, _P, _P, _D, _P]
        L.ora_cons_to_prim.argtypes = [_D, _P, _D, _P]
        L.ora_prim_to_cons.argtypes = [_D, _P, _D, _P]
        L.ora_lr_states.argtypes = [_D, _I, _P, _P, _D, _D, _I, _I, _P, _P]
        _ora = L
    return _ora


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleGrid:
    """The oracle restatement on one grid (CPU, OpenMP)."""

    def __init__(self, d):
        L = oracle_lib()
        bc = d.get("bc", (4, 4, 4, 4))
        flux = d.get("flux", "roe")
        fl = {"roe": 0, "hlle": 1}[flux] if isinstance(flux, str) else int(flux)
        self.h = L.ora_grid_new(int(d["nx1"]), int(d["nx2"]), d["x1min"], d["x1max"], d["x2min"],
                                d["x2max"], d["gamma"], d.get("cour_no", 0.4), *[int(b) for b in bc], fl)
        self.sx, self.sy = int(d["nx1"]) + 8, int(d["nx2"]) + 8
        self.L = L

    def __del__(self):
        try:
            self.L.ora_grid_free(self.h)
        except Exception:
            pass

    def info(self):
        iv = (ctypes.c_int * 9)()
        dv = (ctypes.c_double * 6)()
        self.L.ora_info(self.h, iv, dv)
        return list(iv), list(dv)

    def clock(self):
        iv, dv = self.info()
        return dv[4], dv[5], iv[6]

    def get(self):
        gas = np.empty((self.sy, self.sx, 8))
        b = [np.empty((self.sy, self.sx)) for _ in range(3)]
        self.L.ora_get(self.h, p(gas), *[p(x) for x in b])
        return gas, b[0], b[1], b[2]

    def put(self, gas, b1, b2, b3):
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (gas, b1, b2, b3)]
        self.L.ora_put(self.h, *[p(x) for x in arrs])

    def set_time(self, dt, time, nstep):
        self.L.ora_set_time(self.h, dt, time, nstep)

    def start(self):
        self.L.ora_start(self.h)

    def step(self, tlim):
        self.L.ora_step(self.h, tlim)

    def run(self, tlim, nlim):
        return self.L.ora_run(self.h, tlim, nlim)

    def integrate(self):
        self.L.ora_integrate(self.h)

    def new_dt(self):
        self.L.ora_new_dt(self.h)

    def set_bvals(self):
        self.L.ora_set_bvals(self.h)


def oracle_setup(problem, overrides=(), flux="roe"):
    """main.c:64-151 with the oracle: returns (grid, tlim, nlim)."""
    pkg = load_pkg()
    par = pkg.read_athinput(os.path.join(INPUTS, f"athinput.{problem}"), overrides)
    d = pkg.desc_from_athinput(par, flux=flux)
    g = OracleGrid(d)
    pr = par["problem"]
    f = pkg._atof
    if problem == "blast":
        g.L.ora_pgen_blast(g.h, f(pr["radius"]), f(pr["pamb"]), f(pr["prat"]), f(pr["b0"]),
                           f(pr["angle"]))
    elif problem == "field_loop":
        g.L.ora_pgen_field_loop(g.h, f(pr["rad"]), f(pr["amp"]), f(pr["vflow"]),
                                int(f(pr["iprob"])))
    else:
        g.L.ora_pgen_cpaw2d(g.h, f(pr["angle"]), f(pr["b_par"]), f(pr["b_perp"]),
                            f(pr["pres"]), f(pr["v_par"]))
    g.start()
    return g, f(par["time"]["tlim"]), pkg._atoi(par["time"].get("nlim", "-1"))


def l1_rel(a, b):
    den = np.sum(np.abs(b))
    return float(np.sum(np.abs(a - b)) / den) if den > 0 else float(np.sum(np.abs(a - b)))


def interior(a, nx1, nx2):
    return a[4:4 + nx2, 4:4 + nx1]
