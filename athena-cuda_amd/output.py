"""The reference's output control on a device-resident grid: the <outputN>
blocks of an athinput (init_output, output.c:141-356), the cadence check at
the top of every main-loop iteration (data_output, output.c:362-441) and the
main loop that calls it (main.c:221-318). The files themselves (.bin, .tab,
.ppm) are written by the library (include/athmi_output.h) from the device
state, byte-identical to the reference's.

    g, par, tlim, nlim = setup_problem("blast", [...])
    outs = init_output(par, "blast")
    run_main_loop(g, outs, tlim, nlim, basename=par["job"]["problem_id"], outdir=".")
"""
import ctypes
import math
import os
import sys

from . import AthmiError, _atof, _atoi, _check, lib

MAXOUT_DEFAULT = 10   # output.c:99
NUM_DIGIT = 4         # defs.h:83, digits of the dump number in file names

# output.c:726-768 (MHD, adiabatic build: defs.h:29-32)
BUILTIN_EXPRS = ("d", "M1", "M2", "M3", "E", "B1c", "B2c", "B3c", "ME", "V1", "V2", "V3", "P",
                 "cs2", "S")
# get_usr_expr of each problem (field_loop.c:174-179, blast.c:93, cpaw2d.c:207)
USER_EXPRS = {"blast": (), "field_loop": ("J3", "Bp2"), "cpaw2d": ()}
PALETTES = ("rainbow", "jh_colors", "idl1", "idl2", "step8", "step32", "heat")


class _PpmSpec(ctypes.Structure):
    _fields_ = [("expr", ctypes.c_char_p), ("palette", ctypes.c_char_p),
                ("sdmin", ctypes.c_int), ("sdmax", ctypes.c_int),
                ("dmin", ctypes.c_double), ("dmax", ctypes.c_double)]


def _bind():
    L = lib()
    if not getattr(L, "_output_bound", False):
        _P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.athmi_dump_binary.argtypes = [_P, _I, ctypes.c_char_p]
        L.athmi_dump_tab.argtypes = [_P, _I, ctypes.c_char_p, ctypes.c_char_p]
        L.athmi_output_ppm.argtypes = [_P, _I, ctypes.c_char_p, ctypes.POINTER(_PpmSpec), _I,
                                       ctypes.POINTER(_D), ctypes.POINTER(_D)]
        for n in ("athmi_dump_binary", "athmi_dump_tab", "athmi_output_ppm"):
            getattr(L, n).restype = _I
        L._output_bound = True
    return L


class Output:
    """One <outputN> block (athena.h:145-183, the members the 2-D path uses)."""

    def __init__(self, n, dt, t, num, out, id_, out_fmt):
        self.n, self.dt, self.t, self.num = n, dt, t, num
        self.out, self.id, self.out_fmt = out, id_, out_fmt
        self.dat_fmt = None
        self.dmin = self.dmax = 0.0
        self.sdmin = self.sdmax = 0
        self.gmin = self.gmax = 0.0
        self.palette = None
        self.ndim = 2

    def __repr__(self):
        return (f"Output(n={self.n}, fmt={self.out_fmt}, out={self.out}, id={self.id}, "
                f"dt={self.dt}, t={self.t}, num={self.num})")


def fname_construct(basename, dlen, idump, id_, ext):
    """<basename>[.idump][.id].<ext> (ath_files.c:31-63)."""
    num = f".{idump:0{dlen}d}" if dlen > 0 and idump >= 0 else ""
    mid = f".{id_}" if id_ is not None else ""
    return f"{basename}{num}{mid}.{ext}"


def _warn(msg):
    print(msg, file=sys.stderr)


def _slice_is_set(block, axname, nx):
    """parse_slice (output.c:800-850): True when <outputN>/ixK selects
    indices (the output then loses that dimension)."""
    if axname not in block:
        return False
    expr = block[axname]
    if ":" in expr:
        lo_s = expr.split(":", 1)[0].strip()
        lo = _atoi(lo_s) - 1 if lo_s else 0
    else:
        lo = _atoi(expr) - 1
    if lo < 0 or lo >= nx:
        raise AthmiError(f"[parse_slice]: {expr} has illegal slice {lo}, not in range 1..{nx}")
    return True


def init_output(par, problem, time=0.0):
    """output.c:141-356: the outputs of the first job/maxout <outputN> blocks."""
    maxout = _atoi(par.get("job", {}).get("maxout", str(MAXOUT_DEFAULT)))
    grid = par["grid"]
    nx1, nx2 = _atoi(grid["Nx1"]), _atoi(grid["Nx2"])
    outs = []
    for outn in range(1, maxout + 1):
        name = f"output{outn}"
        b = par.get(name)
        if b is None or "out_fmt" not in b:
            continue
        if "dt" not in b:
            raise AthmiError(f"par_getd: {name}/dt does not exist")
        o = Output(outn, _atof(b["dt"]), _atof(b["time"]) if "time" in b else time,
                   _atoi(b.get("num", "0")), b.get("out", "all"), b.get("id", f"out{outn}"),
                   b["out_fmt"])
        fmt = o.out_fmt
        if o.out == "all":
            if fmt == "bin":
                outs.append(o)
                continue
            if fmt == "tab":
                o.dat_fmt = b.get("dat_fmt")
                outs.append(o)
                continue
            raise AthmiError(f"Unsupported dump mode for {name}/out_fmt={fmt} for out=all")
        usr = _atoi(b["usr_expr_flag"]) if "usr_expr_flag" in b else 0
        known = USER_EXPRS.get(problem, ()) if usr else BUILTIN_EXPRS
        if o.out not in known:
            _warn(f"Could not parse expression {o.out}, skipping it")
            continue
        o.ndim = (0 if _slice_is_set(b, "ix1", nx1) else 1) + (0 if _slice_is_set(b, "ix2", nx2) else 1)
        if "dmin" in b:
            o.sdmin, o.dmin = 1, _atof(b["dmin"])
        if "dmax" in b:
            o.sdmax, o.dmax = 1, _atof(b["dmax"])
        if fmt == "ppm":
            o.palette = b.get("palette", "rainbow")
            if o.palette not in PALETTES:
                raise AthmiError(f"could not find palette={o.palette}, valid names are: "
                                 + " ".join(PALETTES))
        else:
            raise AthmiError(f"Unsupported {name}/out_fmt={fmt}")
        outs.append(o)
    return outs


def write_output(g, o, basename, outdir=".", inst=0):
    """The output function of one Output (dump_binary, dump_tab, output_ppm)."""
    L = _bind()
    if o.out_fmt == "bin":
        path = os.path.join(outdir, fname_construct(basename, NUM_DIGIT, o.num, None, "bin"))
        _check(L.athmi_dump_binary(g._h, inst, path.encode()), "athmi_dump_binary")
    elif o.out_fmt == "tab":
        path = os.path.join(outdir, fname_construct(basename, NUM_DIGIT, o.num, None, "tab"))
        fmt = o.dat_fmt.encode() if o.dat_fmt is not None else None
        _check(L.athmi_dump_tab(g._h, inst, path.encode(), fmt), "athmi_dump_tab")
    elif o.out_fmt == "ppm":
        if o.ndim != 2:
            raise AthmiError("[output_ppm:] Data must be 2D")
        path = os.path.join(outdir, fname_construct(basename, NUM_DIGIT, o.num, o.id, "ppm"))
        spec = _PpmSpec(o.out.encode(), o.palette.encode(), o.sdmin, o.sdmax, o.dmin, o.dmax)
        gmin, gmax = ctypes.c_double(o.gmin), ctypes.c_double(o.gmax)
        _check(L.athmi_output_ppm(g._h, inst, path.encode(), ctypes.byref(spec), o.num,
                                  ctypes.byref(gmin), ctypes.byref(gmax)), "athmi_output_ppm")
        o.gmin, o.gmax = gmin.value, gmax.value
    else:
        raise AthmiError(f"Unsupported out_fmt={o.out_fmt}")
    return path


def data_output(g, outputs, flag, basename, outdir=".", inst=0):
    """output.c:362-441: write every output whose time has come (or all, flag
    != 0); a due output's next time advances by one dt. Returns the paths."""
    time = g.clock(inst)[1]
    due = []
    for o in outputs:
        f = flag
        if time >= o.t:
            o.t += o.dt
            f = 1
        due.append(f)
    paths = []
    for o, f in zip(outputs, due):
        if f:
            paths.append(write_output(g, o, basename, outdir, inst))
            o.num += 1
    return paths


def run_main_loop(g, outputs, tlim, nlim, basename, outdir="."):
    """main.c:221-318 with data_output(.., 0) at the top of every iteration.
    Between outputs the steps run on the device without host round trips:
    the clock waits once time reaches the earliest next output time
    (athmi_set_pause). Returns (steps, files written)."""
    _, _, n0 = g.clock()
    files = []
    try:
        while True:
            _, time, nstep = g.clock()
            if not (time < tlim and (nlim < 0 or nstep < nlim)):
                break
            files += data_output(g, outputs, 0, basename, outdir)
            g.set_pause(math.inf)  # this iteration's outputs are written: step once
            g.step()
            if outputs:
                g.set_pause(min(o.t for o in outputs))
            g.run(-1)
    finally:
        g.set_pause(math.inf)
    return g.clock()[2] - n0, files
