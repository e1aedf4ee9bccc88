"""MI355X-native 2-D CTU + constrained-transport MHD step (the Athena 3.1 path
of awasiljew/athena-cuda) -- Python host side over the C-ABI of libathmi.so.

The surface mirrors the reference's: athinput files with block/name=value
overrides (par.c:133-212), a Grid whose state lives in HBM, the problem
generators blast / field_loop / cpaw2d, and the main loop of main.c:221-318
(clamp dt to tlim, integrate_2d, advance, new_dt, set_bvals).

PyTorch is not needed here; the differentiation form (athena_cuda_amd.adjoint)
uses it for device tensors and autograd.
"""
import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libathmi.so")
INPUTS = os.path.join(_HERE, "inputs")

BC_REFLECT, BC_OUTFLOW, BC_PERIODIC = 1, 2, 4
FAIL_NONFINITE, FAIL_DENSITY, FAIL_PRESSURE = 1, 2, 4
FLUX_ROE, FLUX_HLLE = 0, 1
EXACT, FAST = 0, 1
VARIANTS = {"exact": EXACT, "fast": FAST}
FLUXES = {"roe": FLUX_ROE, "hlle": FLUX_HLLE}


class AthmiError(RuntimeError):
    pass


class _Desc(ctypes.Structure):
    _fields_ = [("nx1", ctypes.c_int), ("nx2", ctypes.c_int),
                ("x1min", ctypes.c_double), ("x1max", ctypes.c_double),
                ("x2min", ctypes.c_double), ("x2max", ctypes.c_double),
                ("gamma", ctypes.c_double), ("cour_no", ctypes.c_double),
                ("bc_ix1", ctypes.c_int), ("bc_ox1", ctypes.c_int),
                ("bc_ix2", ctypes.c_int), ("bc_ox2", ctypes.c_int),
                ("flux", ctypes.c_int), ("variant", ctypes.c_int)]


_lib = None

# exported C-ABI: name -> (restype, argtypes)
_D, _I, _P, _DP, _IP = ctypes.c_double, ctypes.c_int, ctypes.c_void_p, \
    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
SIGNATURES = {
    "athmi_create": (_I, [ctypes.POINTER(_Desc), _I, ctypes.POINTER(_P)]),
    "athmi_destroy": (_I, [_P]),
    "athmi_count": (_I, [_P, _IP]),
    "athmi_set_spacing": (_I, [_P, _I, _D, _D]),
    "athmi_upload": (_I, [_P, _I, _P, _P, _P, _P]),
    "athmi_download": (_I, [_P, _I, _P, _P, _P, _P]),
    "athmi_set_clock": (_I, [_P, _I, _D, _D, _I]),
    "athmi_get_clock": (_I, [_P, _I, _DP, _DP, _IP]),
    "athmi_set_limits": (_I, [_P, _I, _D, _I]),
    "athmi_get_failure": (_I, [_P, _I, _IP, _IP]),
    "athmi_set_fail_stop": (_I, [_P, _I, _I]),
    "athmi_set_pause": (_I, [_P, _I, _D]),
    "athmi_set_bvals": (_I, [_P]),
    "athmi_new_dt": (_I, [_P]),
    "athmi_integrate": (_I, [_P]),
    "athmi_step": (_I, [_P]),
    "athmi_run": (_I, [_P, _I, _IP]),
    "athmi_divb_max": (_I, [_P, _I, _DP]),
    "athmi_set_stream": (_I, [_P, _P]),
    "athmi_sync": (_I, [_P]),
    "athmi_set_shape": (_I, [_P, _I, _I]),
    "athmi_use_graph": (_I, [_P, _I]),
    "athmi_autotune": (_I, [_P, _I, _IP, _IP, _IP, _DP]),
    "athmi_set_tile_order": (_I, [_P, _I]),
    "athmi_set_corr_rows": (_I, [_P, _I]),
    "athmi_set_x2_kind": (_I, [_P, _I]),
    "athmi_get_x2_kind": (_I, [_P, _IP]),
    "athmi_set_fusion": (_I, [_P, _I]),
    "athmi_get_fusion": (_I, [_P, _IP]),
    "athmi_set_strip": (_I, [_P, _I]),
    "athmi_get_strip": (_I, [_P, _IP]),
    "athmi_set_final_rows": (_I, [_P, _I]),
    "athmi_get_final_rows": (_I, [_P, _IP]),
    "athmi_get_shape": (_I, [_P, _IP, _IP, _IP, _IP]),
    "athmi_device_state": (_I, [_P, _I, ctypes.POINTER(_P), _IP]),
    "athmi_pgen_blast": (_I, [_P, _I, _D, _D, _D, _D, _D, _D, _D]),
    "athmi_pgen_field_loop": (_I, [_P, _I, _D, _D, _D, _I]),
    "athmi_pgen_cpaw2d": (_I, [_P, _I, _D, _D, _D, _D, _D, _DP]),
    "athmi_cpaw2d_l1": (_I, [_P, _I, _DP]),
    "athmi_pgen_device": (_I, [_P, _I, _I, _P, _DP]),
    "athmi_integrate_static": (_I, [ctypes.POINTER(_Desc), _D, _P, _P, _I, ctypes.c_size_t,
                                    _P]),
    "athmi_last_error": (ctypes.c_char_p, []),
}


def lib():
    """Load libathmi.so (built in-tree by athena-cuda_amd/build.py). There is
    no CPU fallback: a missing library is an error."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AthmiError(f"{LIB_PATH} is missing: build it with "
                             "python athena-cuda_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise AthmiError(f"{what} failed ({rc}): {lib().athmi_last_error().decode()}")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------------------
# athinput (par.c:133-212): <block> headers, name = value # comment lines,
# <par_end>; block/name=value overrides must name an existing entry.
def read_athinput(path, overrides=()):
    blocks, cur = {}, None
    with open(path) as fh:
        for raw in fh:
            line = raw.strip()
            if not line or line.startswith("#"):
                continue
            if line.startswith("<par_end>"):
                break
            if line.startswith("<"):
                cur = line[1:line.index(">")].strip()
                blocks.setdefault(cur, {})
                continue
            if cur is None or "=" not in line:
                continue
            name, val = line.split("=", 1)
            val = val.split("#", 1)[0].strip()
            blocks[cur][name.strip()] = val
    for ov in overrides:
        m = re.fullmatch(r"([^/]+)/([^=]+)=(.*)", ov)
        if not m:
            continue
        b, n, v = m.groups()
        if b not in blocks:
            raise AthmiError(f'par_cmdline: Block "{b}" not found')
        if n not in blocks[b]:
            raise AthmiError(f'par_cmdline: Par "{n}" not found')
        blocks[b][n] = v
    return blocks


def _atof(s):
    m = re.match(r"\s*[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?", s)
    return float(m.group(0)) if m else 0.0


def _atoi(s):
    m = re.match(r"\s*[-+]?\d+", s)
    return int(m.group(0)) if m else 0


def desc_from_athinput(par, flux="roe", variant="exact"):
    g = par["grid"]
    return dict(nx1=_atoi(g["Nx1"]), nx2=_atoi(g["Nx2"]),
                x1min=_atof(g["x1min"]), x1max=_atof(g["x1max"]),
                x2min=_atof(g["x2min"]), x2max=_atof(g["x2max"]),
                gamma=_atof(par["problem"]["gamma"]), cour_no=_atof(par["time"]["cour_no"]),
                bc=(_atoi(g["ibc_x1"]), _atoi(g["obc_x1"]), _atoi(g["ibc_x2"]), _atoi(g["obc_x2"])),
                flux=flux, variant=variant)


class _PgenParams(ctypes.Structure):
    _fields_ = [("problem", ctypes.c_int), ("p", ctypes.c_double * 7)]


class Grid:
    """One or more independent grid instances resident in HBM (a plain grid
    is a batch of one). descs: list of dicts with nx1, nx2, x1min, x1max,
    x2min, x2max, gamma, cour_no, bc (4 flags), flux ('roe'|'hlle')."""

    def __init__(self, descs, variant="exact"):
        if isinstance(descs, dict):
            descs = [descs]
        arr = (_Desc * len(descs))()
        for k, d in enumerate(descs):
            bc = d.get("bc", (4, 4, 4, 4))
            flux = d.get("flux", "roe")
            arr[k] = _Desc(int(d["nx1"]), int(d["nx2"]), d["x1min"], d["x1max"], d["x2min"],
                           d["x2max"], d["gamma"], d.get("cour_no", 0.4), *[int(b) for b in bc],
                           FLUXES[flux] if isinstance(flux, str) else int(flux),
                           VARIANTS[variant] if isinstance(variant, str) else int(variant))
        h = ctypes.c_void_p()
        _check(lib().athmi_create(arr, len(descs), ctypes.byref(h)), "athmi_create")
        self._h = h
        self.descs = [dict(d) for d in descs]
        self.n = len(descs)

    def close(self):
        if getattr(self, "_h", None):
            lib().athmi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, inst=0):
        d = self.descs[inst]
        return int(d["nx2"]) + 8, int(d["nx1"]) + 8

    # --- state -----------------------------------------------------------
    def upload(self, inst, gas, b1i, b2i, b3i):
        sy, sx = self.shape(inst)
        gas = np.ascontiguousarray(gas, dtype=np.float64).reshape(sy, sx, 8)
        b = [np.ascontiguousarray(x, dtype=np.float64).reshape(sy, sx) for x in (b1i, b2i, b3i)]
        _check(lib().athmi_upload(self._h, inst, _ptr(gas), *[_ptr(x) for x in b]), "upload")

    def download(self, inst=0):
        sy, sx = self.shape(inst)
        gas = np.empty((sy, sx, 8))
        b = [np.empty((sy, sx)) for _ in range(3)]
        _check(lib().athmi_download(self._h, inst, _ptr(gas), *[_ptr(x) for x in b]), "download")
        return gas, b[0], b[1], b[2]

    def clock(self, inst=0):
        dt, t, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        _check(lib().athmi_get_clock(self._h, inst, ctypes.byref(dt), ctypes.byref(t),
                                     ctypes.byref(n)), "get_clock")
        return dt.value, t.value, n.value

    def set_clock(self, inst, dt, time, nstep):
        _check(lib().athmi_set_clock(self._h, inst, dt, time, nstep), "set_clock")

    def failure(self, inst=0):
        """(flags, nstep): the FAIL_* bits raised by the states this instance
        produced and the step of the first one (-1: none). An instance whose
        flags meet its stop set (set_fail_stop) no longer steps."""
        f, n = ctypes.c_int(), ctypes.c_int()
        _check(lib().athmi_get_failure(self._h, inst, ctypes.byref(f), ctypes.byref(n)),
               "get_failure")
        return f.value, n.value

    def set_fail_stop(self, mask, inst=-1):
        """FAIL_* bits that stop the main-loop step (default FAIL_NONFINITE |
        FAIL_DENSITY; a pressure-floor hit is only reported, as the reference
        floors and steps on)"""
        _check(lib().athmi_set_fail_stop(self._h, inst, int(mask)), "set_fail_stop")

    def set_limits(self, tlim, nlim=-1, inst=-1):
        _check(lib().athmi_set_limits(self._h, inst, tlim, nlim), "set_limits")

    def set_pause(self, t_next, inst=-1):
        """the main-loop step waits once time >= t_next (an output is due)"""
        _check(lib().athmi_set_pause(self._h, inst, t_next), "set_pause")

    # --- the reference's routines -------------------------------------------
    def set_bvals(self):
        _check(lib().athmi_set_bvals(self._h), "set_bvals")

    def new_dt(self):
        _check(lib().athmi_new_dt(self._h), "new_dt")

    def integrate(self):
        _check(lib().athmi_integrate(self._h), "integrate")

    def step(self):
        _check(lib().athmi_step(self._h), "step")

    def run(self, max_steps=-1):
        n = ctypes.c_int()
        _check(lib().athmi_run(self._h, max_steps, ctypes.byref(n)), "run")
        return n.value

    def sync(self):
        _check(lib().athmi_sync(self._h), "sync")

    def divb_max(self, inst=0):
        v = ctypes.c_double()
        _check(lib().athmi_divb_max(self._h, inst, ctypes.byref(v)), "divb_max")
        return v.value

    def set_shape(self, x2_seg=32, persist=0):
        _check(lib().athmi_set_shape(self._h, x2_seg, persist), "set_shape")

    def autotune(self, steps=2):
        """Pick the fastest launch shape for this batch (state is restored).
        Returns (x2_seg, persist, xcd_remap, ms_per_step)."""
        a, b, c, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().athmi_autotune(self._h, steps, ctypes.byref(a), ctypes.byref(b),
                                    ctypes.byref(c), ctypes.byref(ms)), "autotune")
        return a.value, b.value, c.value, ms.value

    def set_corr_rows(self, rows=4):
        _check(lib().athmi_set_corr_rows(self._h, int(rows)), "set_corr_rows")

    def launch_shape(self):
        """(x2_seg, persist, xcd_remap, corr_rows) of the current launch shape"""
        v = [ctypes.c_int() for _ in range(4)]
        _check(lib().athmi_get_shape(self._h, *[ctypes.byref(x) for x in v]), "get_shape")
        return tuple(x.value for x in v)

    def set_x2_kind(self, kind):
        """x2 predictor form: 0 = stencil in registers, 1 = stencil in LDS"""
        _check(lib().athmi_set_x2_kind(self._h, kind), "set_x2_kind")

    def x2_kind(self):
        k = ctypes.c_int()
        _check(lib().athmi_get_x2_kind(self._h, ctypes.byref(k)), "get_x2_kind")
        return k.value

    def set_fusion(self, depth):
        """0: the corner emf of Steps 4 and 9 in kernels of their own; 1: the Step 9
        corners inside the final kernel; 2: also the Step 4 corners inside the x1
        corrector"""
        _check(lib().athmi_set_fusion(self._h, depth), "set_fusion")

    def fusion(self):
        k = ctypes.c_int()
        _check(lib().athmi_get_fusion(self._h, ctypes.byref(k)), "get_fusion")
        return k.value

    def set_strip(self, tile_cols):
        """Tile order of the corrector and final kernels: 0 row by row, s > 0
        strips of s tile columns"""
        _check(lib().athmi_set_strip(self._h, int(tile_cols)), "set_strip")

    def strip(self):
        k = ctypes.c_int()
        _check(lib().athmi_get_strip(self._h, ctypes.byref(k)), "get_strip")
        return k.value

    def set_final_rows(self, rows):
        """Rows of a final-update tile: 4, 8 or 16"""
        _check(lib().athmi_set_final_rows(self._h, int(rows)), "set_final_rows")

    def final_rows(self):
        k = ctypes.c_int()
        _check(lib().athmi_get_final_rows(self._h, ctypes.byref(k)), "get_final_rows")
        return k.value

    def set_tile_order(self, xcd_remap=1):
        _check(lib().athmi_set_tile_order(self._h, int(xcd_remap)), "set_tile_order")

    def use_graph(self, on=True):
        _check(lib().athmi_use_graph(self._h, 1 if on else 0), "use_graph")

    def set_stream(self, stream_ptr):
        _check(lib().athmi_set_stream(self._h, ctypes.c_void_p(stream_ptr)), "set_stream")

    def device_state(self, inst=0):
        planes = (ctypes.c_void_p * 11)()
        pitch = ctypes.c_int()
        _check(lib().athmi_device_state(self._h, inst, planes, ctypes.byref(pitch)),
               "device_state")
        return [p for p in planes], pitch.value

    # --- problem generators ------------------------------------------------
    def pgen_blast(self, inst=0, radius=0.1, pamb=0.1, prat=100.0, b0=1.0, angle=45.0,
                   cx=0.0, cy=0.0):
        _check(lib().athmi_pgen_blast(self._h, inst, radius, pamb, prat, b0, angle, cx, cy),
               "pgen_blast")

    def pgen_field_loop(self, inst=0, rad=0.3, amp=1e-3, vflow=1.0, iprob=1):
        _check(lib().athmi_pgen_field_loop(self._h, inst, rad, amp, vflow, iprob),
               "pgen_field_loop")

    def pgen_cpaw2d(self, inst=0, angle=63.4, b_par=1.0, b_perp=0.1, pres=0.1, v_par=0.0):
        lam = ctypes.c_double()
        _check(lib().athmi_pgen_cpaw2d(self._h, inst, angle, b_par, b_perp, pres, v_par,
                                       ctypes.byref(lam)), "pgen_cpaw2d")
        return lam.value

    def pgen_device(self, first, problems):
        """the problem generators on the device for instances first,
        first+1, ...: problems is a list of (name, params) with name in
        PROBLEMS and params in the order of the host generators. Returns the
        cpaw2d wavelengths (0 for the other problems)."""
        n = len(problems)
        arr = (_PgenParams * n)()
        for m, (name, prm) in enumerate(problems):
            arr[m].problem = PROBLEMS.index(name)
            for q, v in enumerate(prm):
                arr[m].p[q] = float(v)
        lam = (ctypes.c_double * n)()
        _check(lib().athmi_pgen_device(self._h, first, n, arr, lam), "pgen_device")
        return list(lam)

    def cpaw2d_l1(self, inst=0):
        out = (ctypes.c_double * 9)()
        _check(lib().athmi_cpaw2d_l1(self._h, inst, out), "cpaw2d_l1")
        return list(out)


PROBLEMS = ("blast", "field_loop", "cpaw2d")


def setup_problem(problem, overrides=(), flux="roe", variant="exact", athinput=None):
    """main.c:64-151 on the device: read the athinput (+ overrides), build the
    grid, run the problem generator, set the ghost zones and the first dt.
    Returns (grid, params, tlim, nlim)."""
    path = athinput or os.path.join(INPUTS, f"athinput.{problem}")
    par = read_athinput(path, overrides)
    g = Grid(desc_from_athinput(par, flux=flux), variant=variant)
    pr = par["problem"]
    if problem == "blast":
        g.pgen_blast(0, _atof(pr["radius"]), _atof(pr["pamb"]), _atof(pr["prat"]),
                     _atof(pr["b0"]), _atof(pr["angle"]))
    elif problem == "field_loop":
        g.pgen_field_loop(0, _atof(pr["rad"]), _atof(pr["amp"]), _atof(pr["vflow"]),
                          int(_atof(pr["iprob"])))
    elif problem == "cpaw2d":
        g.pgen_cpaw2d(0, _atof(pr["angle"]), _atof(pr["b_par"]), _atof(pr["b_perp"]),
                      _atof(pr["pres"]), _atof(pr["v_par"]))
    else:
        raise AthmiError(f"unknown problem {problem}")
    g.set_bvals()
    g.new_dt()
    tlim = _atof(par["time"]["tlim"])
    nlim = _atoi(par["time"].get("nlim", "-1"))
    g.set_limits(tlim, nlim)
    return g, par, tlim, nlim
