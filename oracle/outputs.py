"""TEST INFRASTRUCTURE ONLY (the checker, never the product): a numpy
restatement of the reference's output formats on a host state -- dump_binary
(dump_binary.c:25-134), dump_tab (dump_tab.c:29-233), output_ppm with its
expressions, minmax2 and compute_rgb (output_ppm.c:37-180, output.c:558-741,
utils.c:148-162) and the cadence of data_output (output.c:362-441). Only
tests/ use it, to check the device-side outputs.

State arrays are the reference's layouts: gas [Nx2+8][Nx1+8][8] (Gas members
d, M1, M2, M3, E, B1c, B2c, B3c) and the interface fields [Nx2+8][Nx1+8].
Palettes are passed in as float32 [768] arrays.
"""
import numpy as np

NG = 4  # ghost zones


class Geom:
    """init_grid.c:25-118 / cc_pos.c:25-31 for a grid with 4 ghost zones."""

    def __init__(self, nx1, nx2, x1min, x2min, dx1, dx2):
        self.nx1, self.nx2 = nx1, nx2
        self.is_, self.ie, self.js, self.je = NG, nx1 + NG - 1, NG, nx2 + NG - 1
        self.x1_0, self.x2_0, self.dx1, self.dx2 = x1min, x2min, dx1, dx2

    def cc_pos(self, i, j):
        return (self.x1_0 + ((i - self.is_) + 0.5) * self.dx1,
                self.x2_0 + ((j - self.js) + 0.5) * self.dx2)


def fname(basename, num, id_, ext):
    """fname_construct (ath_files.c:31-63) with num_digit = 4."""
    return f"{basename}.{num:04d}" + (f".{id_}" if id_ is not None else "") + f".{ext}"


def dump_binary(path, gas, G, time, dt, gamma_1):
    with open(path, "wb") as f:
        f.write(np.array([G.nx1, G.nx2, 0, 8, 0, 0], dtype=np.int32).tobytes())
        f.write(np.array([gamma_1, 0.0], dtype=np.float32).tobytes())
        f.write(np.array([time, dt], dtype=np.float32).tobytes())
        xs = [G.cc_pos(i, G.js)[0] for i in range(G.is_, G.ie + 1)]
        ys = [G.cc_pos(G.is_, j)[1] for j in range(G.js, G.je + 1)]
        f.write(np.array(xs, dtype=np.float32).tobytes())
        f.write(np.array(ys, dtype=np.float32).tobytes())
        inner = gas[G.js:G.je + 1, G.is_:G.ie + 1, :]
        f.write(np.ascontiguousarray(np.moveaxis(inner, 2, 0)).astype(np.float32).tobytes())


def dump_tab(path, gas, b1i, b2i, b3i, G, time, gamma_1, dat_fmt=None):
    fmt = " " + (dat_fmt if dat_fmt is not None else "%12.8e")
    il, iu, jl, ju = G.is_, G.ie, G.js, G.je
    j = (ju + jl) // 2
    out = []
    if G.nx1 > 1:
        out.append("# Nx1 = %d\n" % (iu - il + 1))
        out.append("# x1-size = %g\n" % ((iu - il + 1) * G.dx1))
    if G.nx2 > 1:
        out.append("# Nx2 = %d\n" % (ju - jl + 1))
        out.append("# x2-size = %g\n" % ((ju - jl + 1) * G.dx2))
    out.append("# Time = %g\n" % time)
    out.append("#\n#")
    col = 1
    heads = []
    if G.nx1 > 1:
        heads.append("i-zone")
    if G.nx2 > 2:
        heads.append("j-zone")
    if G.nx1 > 1:
        heads.append("x1")
    if G.nx2 > 2:
        heads.append("x2")
    heads += ["d", "M1", "M2", "M3", "P", "E", "B1c", "B2c", "B3c", "B1i", "B1i", "B1i"]
    for h in heads:
        out.append(" [%d]=%s" % (col, h))
        col += 1
    out.append("\n#\n")
    for i in range(il, iu + 1):
        x1, x2 = G.cc_pos(i, j)
        if G.nx1 > 1:
            out.append("%d " % i)
        if G.nx2 > 1:
            out.append("%d " % j)
        if G.nx1 > 1:
            out.append(fmt % x1)
        if G.nx2 > 1:
            out.append(fmt % x2)
        d, m1, m2, m3, e, b1, b2, b3 = (float(v) for v in gas[j, i])
        for v in (d, m1, m2, m3):
            out.append(fmt % v)
        ke = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / d
        me = 0.5 * (b1 * b1 + b2 * b2 + b3 * b3)
        out.append(fmt % ((e - ke - me) * gamma_1))
        out.append(fmt % e)
        for v in (b1, b2, b3, b1i[j, i], b2i[j, i], b3i[j, i]):
            out.append(fmt % float(v))
        out.append("\n")
    with open(path, "w") as f:
        f.write("".join(out))


def expression(name, gas, b1i, b2i, G, gamma):
    """subset2 (output.c:558-600) of one expression (output.c:603-741,
    field_loop.c:160-170) as float32 [Nx2][Nx1]."""
    s = (slice(G.js, G.je + 1), slice(G.is_, G.ie + 1))
    U = {n: gas[s + (q,)] for q, n in enumerate(("d", "M1", "M2", "M3", "E", "B1c", "B2c", "B3c"))}
    g1 = gamma - 1.0
    if name in U:
        v = U[name]
    elif name == "ME":
        v = 0.5 * (U["B1c"] * U["B1c"] + U["B2c"] * U["B2c"] + U["B3c"] * U["B3c"])
    elif name in ("V1", "V2", "V3"):
        v = U["M" + name[1]] / U["d"]
    elif name in ("P", "cs2", "S"):
        x = (U["E"] - 0.5 * (U["B1c"] * U["B1c"] + U["B2c"] * U["B2c"] + U["B3c"] * U["B3c"])
             - 0.5 * (U["M1"] * U["M1"] + U["M2"] * U["M2"] + U["M3"] * U["M3"]) / U["d"])
        if name == "P":
            v = g1 * x
        elif name == "cs2":
            v = gamma * g1 * x / U["d"]
        else:
            v = (g1 * x) / np.power(U["d"], gamma)
    elif name == "J3":
        js, ie, is_, je = G.js, G.ie, G.is_, G.je
        v = ((b2i[js:je + 1, is_:ie + 1] - b2i[js:je + 1, is_ - 1:ie]) / G.dx1 -
             (b1i[js:je + 1, is_:ie + 1] - b1i[js - 1:je, is_:ie + 1]) / G.dx2)
    elif name == "Bp2":
        v = U["B1c"] * U["B1c"] + U["B2c"] * U["B2c"]
    else:
        raise KeyError(name)
    return np.ascontiguousarray(v).astype(np.float32)


def minmax2(data):
    """utils.c:148-162: the sequential MIN/MAX scan (defs.h:90-93) -- a NaN
    restarts it, and among equal values the later one stays."""
    flat = data.ravel()
    nan = np.flatnonzero(np.isnan(flat))
    rest = flat[nan[-1] + 1:] if nan.size else flat
    if rest.size == 0:
        return np.float32(np.nan), np.float32(np.nan)
    mn = rest[np.flatnonzero(rest == rest.min())[-1]]
    mx = rest[np.flatnonzero(rest == rest.max())[-1]]
    return np.float32(mn), np.float32(mx)


def compute_rgb(data, dmin, dmax, rgb):
    """output_ppm.c:144-180 over a float32 image, vectorised with the same
    float / double evaluation. Returns uint8 [n][3]."""
    rgb = np.asarray(rgb, dtype=np.float32)
    der = np.zeros(768, dtype=np.float32)
    der[:765] = rgb[3:] - rgb[:765]
    d = np.where(np.isnan(data), 0.0, data.astype(np.float64)).ravel()
    mn, mx = float(dmin), float(dmax)
    out = np.zeros((d.size, 3), dtype=np.int64)
    if mn == mx:
        out[:] = np.where(d > mx, 255, 0)[:, None]
        return out.astype(np.uint8)
    with np.errstate(invalid="ignore", divide="ignore"):
        x = ((d - mn) * 255.0 / (mx - mn)).astype(np.float32)
    lo, hi = x <= 0.0, x >= 255.0
    # a NaN scale (dmin or dmax NaN) indexes the reference's palette out of
    # range; here, as in the device code, such pixels are black
    mid = ~(lo | hi) & ~np.isnan(x)
    for c in range(3):
        edge = np.where(lo, rgb[c], rgb[255 * 3 + c]).astype(np.float64) * 255.0
        i = np.where(mid, x, 0).astype(np.int64)
        w = x - i.astype(np.float32)
        inner = (rgb[i * 3 + c] + w * der[i * 3 + c]).astype(np.float64) * 255.0
        out[:, c] = np.where(mid, np.trunc(inner), np.where(lo | hi, np.trunc(edge), 0.0)).astype(np.int64)
    return out.astype(np.uint8)


def output_ppm(path, data, rgb, out):
    """output_ppm.c:37-137 for one Output-like object `out` (num, sdmin, dmin,
    sdmax, dmax, gmin, gmax are read and gmin/gmax updated)."""
    dmin, dmax = minmax2(data)
    if out.num == 0:
        out.gmin, out.gmax = float(dmin), float(dmax)
    else:
        out.gmin = float(dmin) if float(dmin) < out.gmin else out.gmin
        out.gmax = float(dmax) if float(dmax) > out.gmax else out.gmax
    head = "P6\n# dmin = %.7e, dmax = %.7e, gmin = %.7e, gmax = %.7e\n%d %d\n255\n" % (
        float(dmin), float(dmax), out.gmin, out.gmax, data.shape[1], data.shape[0])
    if out.sdmin:
        dmin = np.float32(out.dmin)
    if out.sdmax:
        dmax = np.float32(out.dmax)
    img = compute_rgb(data[::-1], dmin, dmax, rgb)
    with open(path, "wb") as f:
        f.write(head.encode())
        f.write(img.tobytes())
