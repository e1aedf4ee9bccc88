"""The reference-named drop-in (include/athena_dropin.h) driven the way the
reference's main loop drives its CPU routines (main.c:221-318): integrate_2d,
nstep++ and time += dt, new_dt, set_bvals(Grid*, 0), on a struct with the
layout of Grid (athena.h:106-120) built here through ctypes. The exact variant
must match the oracle value for value, ghosts included, and the returned dt
exactly.

Two Grid layouts: rows in one block, as calloc_2d_array makes them
(ath_array.c:48-66; the drop-in then hands the arrays to the device as they
are), and rows with a stride wider than a row (the drop-in's element-wise
copy)."""
import ctypes

import numpy as np
import pytest

from helpers import INPUTS, load_pkg, oracle_setup

pytestmark = pytest.mark.gpu


class RefGas(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("d", "M1", "M2", "M3", "E", "B1c", "B2c", "B3c")]


PD = ctypes.POINTER(ctypes.c_double)


class RefGrid(ctypes.Structure):
    _fields_ = [("U", ctypes.POINTER(ctypes.POINTER(RefGas))), ("B1i", ctypes.POINTER(PD)),
                ("B2i", ctypes.POINTER(PD)), ("B3i", ctypes.POINTER(PD)),
                ("x1_0", ctypes.c_double), ("x2_0", ctypes.c_double),
                ("dx1", ctypes.c_double), ("dx2", ctypes.c_double),
                ("dt", ctypes.c_double), ("time", ctypes.c_double), ("nstep", ctypes.c_int),
                ("Nx1", ctypes.c_int), ("Nx2", ctypes.c_int), ("is_", ctypes.c_int),
                ("ie", ctypes.c_int), ("js", ctypes.c_int), ("je", ctypes.c_int),
                ("idisp", ctypes.c_int), ("jdisp", ctypes.c_int), ("outfilename", ctypes.c_char_p)]


class HostGrid:
    """Grid arrays in numpy memory; pad > 0 widens each row's storage so the
    rows are not one block."""

    def __init__(self, state, pad):
        gas, b1, b2, b3 = state
        sy, sx = b1.shape
        self.sx, self.sy = sx, sy
        self.gas = np.zeros((sy, sx + pad, 8))
        self.b = [np.zeros((sy, sx + pad)) for _ in range(3)]
        self.set(state)
        self.urows = (ctypes.POINTER(RefGas) * sy)(
            *[ctypes.cast(self.gas[j].ctypes.data, ctypes.POINTER(RefGas)) for j in range(sy)])
        self.brows = [(PD * sy)(*[ctypes.cast(a[j].ctypes.data, PD) for j in range(sy)])
                      for a in self.b]

    def set(self, state):
        gas, b1, b2, b3 = state
        self.gas[:, :self.sx] = np.asarray(gas).reshape(self.sy, self.sx, 8)
        for a, x in zip(self.b, (b1, b2, b3)):
            a[:, :self.sx] = x

    def get(self):
        return (self.gas[:, :self.sx].copy(), *[a[:, :self.sx].copy() for a in self.b])


def check_equal(host, ora, tag):
    for name, x, y in zip(("gas", "B1i", "B2i", "B3i"), host.get(), ora.get()):
        y = np.asarray(y).reshape(x.shape)
        if not np.array_equal(x, y):
            d = np.abs(x - y)
            raise AssertionError(f"{tag}: {name} differs, max |diff| {d.max():.3e} at "
                                 f"{np.unravel_index(np.argmax(d), d.shape)}")


@pytest.mark.parametrize("pad", [0, 3])
@pytest.mark.parametrize("problem,overrides,flux", [
    ("blast", ["grid/Nx1=40", "grid/Nx2=56"], "roe"),
    ("field_loop", ["grid/Nx1=48", "grid/Nx2=32"], "hlle"),
    # thinner than the ghost layers, reflecting: set_bvals reads the ghosts
    # the previous call left
    ("blast", ["grid/Nx1=3", "grid/Nx2=2", "problem/radius=0.2", "grid/ibc_x1=1",
               "grid/obc_x1=1", "grid/ibc_x2=1", "grid/obc_x2=1"], "roe"),
])
def test_dropin_main_loop_matches_oracle(problem, overrides, flux, pad):
    pkg = load_pkg()
    print(f"[gpu] drop-in {problem} {flux} {overrides} pad={pad}", flush=True)
    L = pkg.lib()
    par = pkg.read_athinput(f"{INPUTS}/athinput.{problem}", overrides)
    d = pkg.desc_from_athinput(par, flux=flux)
    ora, tlim, _ = oracle_setup(problem, overrides, flux=flux)
    dt0, t0, n0 = ora.clock()

    host = HostGrid(ora.get(), pad)
    G = RefGrid()
    G.U = host.urows
    G.B1i, G.B2i, G.B3i = host.brows
    G.Nx1, G.Nx2 = d["nx1"], d["nx2"]
    G.x1_0, G.x2_0 = d["x1min"], d["x2min"]
    G.dx1 = (d["x1max"] - d["x1min"]) / d["nx1"]
    G.dx2 = (d["x2max"] - d["x2min"]) / d["nx2"]
    G.is_, G.ie = 4, 4 + d["nx1"] - 1
    G.js, G.je = 4, 4 + d["nx2"] - 1
    G.dt, G.time, G.nstep = dt0, t0, n0

    L.athmi_dropin_configure.argtypes = [ctypes.c_double, ctypes.c_double] + [ctypes.c_int] * 6
    L.athmi_dropin_configure.restype = None
    for name in ("integrate_2d", "new_dt", "set_bvals_init", "integrate_destruct_2d"):
        getattr(L, name).restype = None
    L.set_bvals.argtypes = [ctypes.POINTER(RefGrid), ctypes.c_int]
    L.set_bvals.restype = None
    L.integrate_init_2d.argtypes = [ctypes.c_int, ctypes.c_int]
    L.integrate_init_2d.restype = None

    L.athmi_dropin_configure(d["gamma"], d["cour_no"], *d["bc"], pkg.FLUXES[flux],
                             pkg.VARIANTS["exact"])
    L.integrate_init_2d(G.Nx1, G.Nx2)
    L.set_bvals_init(ctypes.byref(G))
    try:
        for s in range(6):
            L.integrate_2d(ctypes.byref(G))
            ora.integrate()
            check_equal(host, ora, f"integrate_2d #{s}")
            G.nstep += 1
            G.time += G.dt
            ora.set_time(G.dt, G.time, G.nstep)
            L.new_dt(ctypes.byref(G))
            ora.new_dt()
            assert G.dt == ora.clock()[0], f"new_dt #{s}: {G.dt} vs {ora.clock()[0]}"
            L.set_bvals(ctypes.byref(G), 0)
            ora.set_bvals()
            check_equal(host, ora, f"set_bvals #{s}")
    finally:
        L.integrate_destruct_2d()


def bind_dropin(L):
    L.athmi_dropin_configure.argtypes = [ctypes.c_double, ctypes.c_double] + [ctypes.c_int] * 6
    L.athmi_dropin_configure.restype = None
    for name in ("integrate_2d", "new_dt", "set_bvals_init", "integrate_destruct_2d",
                 "athmi_dropin_sync_host", "athmi_dropin_mark_dirty"):
        getattr(L, name).restype = None
    for name in ("integrate_2d", "new_dt", "set_bvals_init", "athmi_dropin_sync_host"):
        getattr(L, name).argtypes = [ctypes.POINTER(RefGrid)]
    L.set_bvals.argtypes = [ctypes.POINTER(RefGrid), ctypes.c_int]
    L.set_bvals.restype = None
    L.athmi_dropin_resident.argtypes = [ctypes.POINTER(RefGrid), ctypes.c_int]
    L.athmi_dropin_resident.restype = None
    L.integrate_init_2d.argtypes = [ctypes.c_int, ctypes.c_int]
    L.integrate_init_2d.restype = None


def ref_grid(host, d, clock):
    G = RefGrid()
    G.U = host.urows
    G.B1i, G.B2i, G.B3i = host.brows
    G.Nx1, G.Nx2 = d["nx1"], d["nx2"]
    G.x1_0, G.x2_0 = d["x1min"], d["x2min"]
    G.dx1 = (d["x1max"] - d["x1min"]) / d["nx1"]
    G.dx2 = (d["x2max"] - d["x2min"]) / d["nx2"]
    G.is_, G.ie = 4, 4 + d["nx1"] - 1
    G.js, G.je = 4, 4 + d["nx2"] - 1
    G.dt, G.time, G.nstep = clock
    return G


@pytest.mark.parametrize("overrides", [
    ["grid/Nx1=64", "grid/Nx2=48"],
    ["grid/Nx1=5", "grid/Nx2=3", "problem/radius=0.2", "grid/ibc_x1=1", "grid/obc_x1=1",
     "grid/ibc_x2=1", "grid/obc_x2=1"]], ids=["blast64x48", "blast5x3_reflect"])
def test_dropin_resident_main_loop_matches_oracle(overrides):
    """Resident mode: the Grid crosses PCIe once; the host arrays are
    refreshed on request (athmi_dropin_sync_host) and equal the oracle's state
    there; a host-side change announced by athmi_dropin_mark_dirty is
    uploaded by the next call. On a reflecting grid thinner than the ghost
    layers the device's set_bvals reads the ghosts its previous call left."""
    pkg = load_pkg()
    L = pkg.lib()
    bind_dropin(L)
    problem, flux = "blast", "roe"
    print(f"[gpu] drop-in resident {problem} {overrides}", flush=True)
    par = pkg.read_athinput(f"{INPUTS}/athinput.{problem}", overrides)
    d = pkg.desc_from_athinput(par, flux=flux)
    ora, tlim, _ = oracle_setup(problem, overrides, flux=flux)
    host = HostGrid(ora.get(), 0)
    G = ref_grid(host, d, ora.clock())
    L.athmi_dropin_configure(d["gamma"], d["cour_no"], *d["bc"], pkg.FLUXES[flux],
                             pkg.VARIANTS["exact"])
    L.integrate_init_2d(G.Nx1, G.Nx2)
    L.set_bvals_init(ctypes.byref(G))
    L.athmi_dropin_resident(ctypes.byref(G), 1)
    before = host.get()
    try:
        for s in range(12):
            L.integrate_2d(ctypes.byref(G))
            ora.integrate()
            G.nstep += 1
            G.time += G.dt
            ora.set_time(G.dt, G.time, G.nstep)
            L.new_dt(ctypes.byref(G))
            ora.new_dt()
            assert G.dt == ora.clock()[0], f"new_dt #{s}"
            L.set_bvals(ctypes.byref(G), 0)
            ora.set_bvals()
            if s == 0:
                # the host arrays lag the device until they are synced
                for x, y in zip(host.get(), before):
                    assert np.array_equal(x, y)
            if s in (4, 11):
                L.athmi_dropin_sync_host(ctypes.byref(G))
                check_equal(host, ora, f"resident step {s}")
            if s == 4:
                # a host-side edit of the state: both sides take it
                gas, b1, b2, b3 = host.get()
                gas[4:8, 4:6, 4] *= 1.5
                host.set((gas, b1, b2, b3))
                ora.put(gas, b1, b2, b3)
                L.athmi_dropin_mark_dirty()
        L.athmi_dropin_resident(ctypes.byref(G), 0)  # syncs the host arrays
        check_equal(host, ora, "after leaving resident mode")
    finally:
        L.integrate_destruct_2d()


def test_dropin_resident_loop_keeps_pace_with_the_device_step():
    """At the benchmark size (blast 4096^2) the reference-named main loop in
    resident mode runs within 2x of athmi_step (the whole iteration replayed
    from the step graph), against about 20x slower with the Grid copied both
    ways on every call."""
    import time
    pkg = load_pkg()
    L = pkg.lib()
    bind_dropin(L)
    ov = ["grid/Nx1=4096", "grid/Nx2=4096"]
    print("[gpu] drop-in resident loop at 4096^2 vs athmi_step", flush=True)
    g, par, tlim, _ = pkg.setup_problem("blast", ov)
    g.set_limits(1e30, -1)
    state0, clock0 = g.download(0), g.clock(0)
    n = 10
    for _ in range(2):
        g.step()
    g.sync()
    t0 = time.perf_counter()
    for _ in range(n):
        g.step()
    g.sync()
    t_step = (time.perf_counter() - t0) / n
    d = g.descs[0]
    g.close()
    host = HostGrid(state0, 0)
    del state0
    G = ref_grid(host, d, clock0)
    L.athmi_dropin_configure(d["gamma"], d["cour_no"], *d["bc"], pkg.FLUXES["roe"],
                             pkg.VARIANTS["exact"])
    L.integrate_init_2d(G.Nx1, G.Nx2)
    L.set_bvals_init(ctypes.byref(G))
    L.athmi_dropin_resident(ctypes.byref(G), 1)

    def iteration():
        if (tlim - G.time) < G.dt:
            G.dt = tlim - G.time
        L.integrate_2d(ctypes.byref(G))
        G.nstep += 1
        G.time += G.dt
        L.new_dt(ctypes.byref(G))
        L.set_bvals(ctypes.byref(G), 0)

    try:
        for _ in range(2):
            iteration()
        t0 = time.perf_counter()
        for _ in range(n):
            iteration()
        t_loop = (time.perf_counter() - t0) / n
    finally:
        L.integrate_destruct_2d()
    print(f"[gpu] athmi_step {1e3 * t_step:.3f} ms, resident drop-in loop "
          f"{1e3 * t_loop:.3f} ms per iteration ({t_loop / t_step:.2f}x)", flush=True)
    assert t_loop <= 2.0 * t_step
