"""athmi_integrate_static (the entry point that allocates nothing: its
descriptor, clock and scratch planes live in a static device array) against
athmi_integrate on a handle, bit for bit, on two calls in a row -- the second
on a smaller grid, so it runs on the scratch the first call left behind --
and its rejection of a grid larger than the static workspace."""
import ctypes

import numpy as np
import pytest
import torch  # before libathmi.so: torch's HIP runtime serves the process

from helpers import load_pkg

pytestmark = pytest.mark.gpu

FIELDS = ("d", "M1", "M2", "M3", "E", "B1c", "B2c", "B3c", "B1i", "B2i", "B3i")


def desc_of(pkg, d, variant="exact"):
    return pkg._Desc(int(d["nx1"]), int(d["nx2"]), d["x1min"], d["x1max"], d["x2min"],
                     d["x2max"], d["gamma"], d.get("cour_no", 0.4), *[int(b) for b in d["bc"]],
                     pkg.FLUXES[d["flux"]], pkg.VARIANTS[variant])


def static_vs_handle(pkg, problem, overrides, flux):
    L = pkg.lib()
    g, par, tlim, _ = pkg.setup_problem(problem, overrides, flux=flux)
    g.set_limits(tlim, 3)  # a few main-loop steps: not the initial state
    g.run()
    dt = g.clock(0)[0]
    planes, pitch = g.device_state(0)
    ps = (planes[1] - planes[0]) // 8
    d = g.descs[0]
    sy, sx = int(d["nx2"]) + 8, int(d["nx1"]) + 8
    # poisoned output: every cell of [sy][sx] must be written (interior
    # integrated, ghosts carried)
    out = torch.full((11 * ps,), float("nan"), dtype=torch.float64, device="cuda:0")
    desc = desc_of(pkg, d)
    print(f"[no_alloc] {problem} {d['nx1']}x{d['nx2']} {flux}: static entry point", flush=True)
    torch.cuda.synchronize()
    rc = L.athmi_integrate_static(ctypes.byref(desc), dt, ctypes.c_void_p(planes[0]),
                                  ctypes.c_void_p(out.data_ptr()), pitch, ps, None)
    pkg._check(rc, "athmi_integrate_static")
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(11, ps)[:, :sy * pitch].reshape(11, sy, pitch)[:, :, :sx]
    g.integrate()
    g.sync()
    gas, b1i, b2i, b3i = g.download(0)
    want = [gas[..., f] for f in range(8)] + [b1i, b2i, b3i]
    for f, name in enumerate(FIELDS):
        assert np.array_equal(got[f], want[f]), \
            f"{problem} {flux}: {name} differs from athmi_integrate " \
            f"(max {np.nanmax(np.abs(got[f] - want[f]))})"
    g.close()


def test_static_entry_point_equals_integrate_twice_second_smaller():
    pkg = load_pkg()
    static_vs_handle(pkg, "blast", ["grid/Nx1=96", "grid/Nx2=144"], "roe")
    static_vs_handle(pkg, "field_loop", ["grid/Nx1=64", "grid/Nx2=32"], "hlle")


def test_static_entry_point_rejects_a_grid_beyond_the_workspace():
    pkg = load_pkg()
    L = pkg.lib()
    d = dict(nx1=1024, nx2=1024, x1min=0.0, x1max=1.0, x2min=0.0, x2max=1.0, gamma=5 / 3,
             bc=(4, 4, 4, 4), flux="roe")
    desc = desc_of(pkg, d)
    pitch = 1040
    out = torch.empty((16,), dtype=torch.float64, device="cuda:0")
    rc = L.athmi_integrate_static(ctypes.byref(desc), 1e-3, ctypes.c_void_p(out.data_ptr()),
                                  ctypes.c_void_p(out.data_ptr()), pitch, pitch * 1032, None)
    assert rc == -1
    assert b"static workspace" in L.athmi_last_error()
