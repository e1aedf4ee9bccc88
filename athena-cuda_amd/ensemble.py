"""Batched uneven ensemble: many independent MHD blast problems of different
grid sizes advanced together by one handle (one launch per pipeline stage over
the tile work-list of all instances, one HIP graph per main-loop step, a dt per
instance).

Each instance is the blast of blast.c:25-72 with the reference's constants
(rho = 1, v = 0, uniform B at an angle, over-pressure inside a radius), drawn
from a seed:

    nx, ny        uniform in {64, 96, ..., 1024}, independently
    domain        x1 in [-0.5, 0.5], x2 in [-ny/(2 nx), ny/(2 nx)] (square cells)
    centre        uniform in the inner 60% of the domain (the reference centres
                  the blast at the origin, blast.c:64; an offset centre is an
                  extension)
    radius        U(0.05, 0.15)
    pressure      ambient 0.1, interior 0.1 * prat with prat log-uniform in [1, 100]
    field         |B| U(0, 1.5) at an angle U(0, pi)
    flux          Roe for even instance indices, HLLE for odd
    gamma, CFL    5/3, 0.4; periodic boundaries

Every instance keeps its own CFL dt (new_dt.c:25-103 per instance) and runs a
fixed number of main-loop steps (nlim).
"""
import math

import numpy as np

from . import Grid

GAMMA = 1.6666666666666667
PAMB = 0.1


def ensemble_params(seed=0, n=256):
    """Per-instance parameters of the ensemble drawn from `seed`."""
    rng = np.random.default_rng(seed)
    sizes = np.arange(64, 1025, 32)
    out = []
    for k in range(n):
        nx = int(rng.choice(sizes))
        ny = int(rng.choice(sizes))
        ly = ny / nx
        cx = float(rng.uniform(-0.3, 0.3))
        cy = float(rng.uniform(-0.3, 0.3)) * ly
        radius = float(rng.uniform(0.05, 0.15))
        prat = float(math.exp(rng.uniform(0.0, math.log(100.0))))
        b0 = float(rng.uniform(0.0, 1.5))
        angle = float(rng.uniform(0.0, math.pi)) * 180.0 / math.pi
        out.append(dict(nx1=nx, nx2=ny, x1min=-0.5, x1max=0.5, x2min=-0.5 * ly, x2max=0.5 * ly,
                        gamma=GAMMA, cour_no=0.4, bc=(4, 4, 4, 4),
                        flux="roe" if k % 2 == 0 else "hlle",
                        cx=cx, cy=cy, radius=radius, prat=prat, b0=b0, angle=angle))
    return out


def cells(params):
    return sum(p["nx1"] * p["nx2"] for p in params)


def build(params, nlim, variant="exact", device_ics=False):
    """One handle holding every instance, initialised as main.c:121-151
    (problem generator, set_bvals, first dt) with main-loop limit nlim.
    device_ics: generate the blasts on the device in one launch (the same
    states as the host generator)."""
    g = Grid(params, variant=variant)
    if device_ics:
        g.pgen_device(0, [("blast", (p["radius"], PAMB, p["prat"], p["b0"], p["angle"], p["cx"],
                                     p["cy"])) for p in params])
    else:
        for k, p in enumerate(params):
            g.pgen_blast(k, p["radius"], PAMB, p["prat"], p["b0"], p["angle"], p["cx"], p["cy"])
    g.set_bvals()
    g.new_dt()
    g.set_limits(1e30, nlim)
    return g
