"""The batched uneven ensemble on the GPU: every instance of a batch must
equal the plain single-problem kernel run on its own (same kernels, same
arithmetic: value for value, clocks included), whatever the launch shape, and
small instances must equal the oracle."""
import importlib

import numpy as np
import pytest

from helpers import OracleGrid, load_pkg

pytestmark = pytest.mark.gpu


def ens():
    pkg = load_pkg()
    return pkg, importlib.import_module(pkg.__name__ + ".ensemble")


def assert_same(a, b, tag):
    for name, x, y in zip(("gas", "b1i", "b2i", "b3i"), a, b):
        assert np.array_equal(x, y), f"{tag} {name}: max diff {np.abs(x - y).max()}"


@pytest.mark.parametrize("shape", [None, (64, 2, 0, 4), (16, 0, 1, 1), (32, 4, 1, 2), (1, 0, 1, 2),
                                   (3, 8, 1, 1)])
def test_batch_instances_equal_plain_runs(shape):
    pkg, E = ens()
    params = E.ensemble_params(seed=7, n=10)
    steps = 40
    print(f"[gpu] ensemble seed 7 x10, {steps} steps, shape {shape}", flush=True)
    gb = E.build(params, nlim=steps)
    if shape:
        gb.set_shape(*shape[:2])
        gb.set_tile_order(shape[2])
        gb.set_corr_rows(shape[3])
        assert gb.launch_shape() == shape
    gb.run()  # issues whole chunks of steps; finished instances idle
    assert all(gb.clock(k)[2] == steps for k in range(gb.n))
    for k, p in enumerate(params):
        g1 = E.build([p], nlim=steps)
        g1.run()
        assert_same(gb.download(k), g1.download(0), f"instance {k} ({p['nx1']}x{p['nx2']})")
        assert gb.clock(k) == g1.clock(0)
        g1.close()
    gb.close()


def test_batch_mixing_boundaries_and_solvers_equals_plain_runs():
    """Instances of one batch with different boundary kinds (set_bvals.c:26-29,
    per instance in the composed ghost map) and Riemann solvers (each predictor
    and corrector launch takes only the tiles of its solver): every instance
    must equal its own plain run, value for value, clocks included."""
    pkg, E = ens()
    params = E.ensemble_params(seed=5, n=6)
    bcs = [(4, 4, 4, 4), (2, 2, 2, 2), (1, 1, 1, 1), (1, 2, 4, 4), (4, 4, 1, 2), (2, 1, 2, 1)]
    for k, p in enumerate(params):
        p["bc"] = bcs[k]
        p["flux"] = "hlle" if k % 3 == 1 else "roe"
    steps = 30
    print(f"[gpu] mixed boundaries and solvers, {steps} steps", flush=True)
    gb = E.build(params, nlim=steps)
    gb.run()
    for k, p in enumerate(params):
        g1 = E.build([p], nlim=steps)
        g1.run()
        assert_same(gb.download(k), g1.download(0), f"instance {k} bc {p['bc']} {p['flux']}")
        assert gb.clock(k) == g1.clock(0)
        g1.close()
    gb.close()


def test_small_ensemble_matches_oracle():
    pkg, E = ens()
    params = E.ensemble_params(seed=3, n=6)
    for k, p in enumerate(params):  # small grids keep the oracle quick
        frac = p["cy"] / (p["x2max"] - p["x2min"])
        p["nx1"], p["nx2"] = 64 + 32 * (k % 2), 64 + 32 * (k // 2 % 2)
        ly = p["nx2"] / p["nx1"]
        p["x2min"], p["x2max"], p["cy"] = -0.5 * ly, 0.5 * ly, frac * ly
    steps = 25
    print(f"[gpu] small ensemble vs oracle, {steps} steps", flush=True)
    gb = E.build(params, nlim=steps)
    gb.run()
    for k, p in enumerate(params):
        o = OracleGrid(p)
        o.L.ora_pgen_blast_at(o.h, p["radius"], E.PAMB, p["prat"], p["b0"], p["angle"], p["cx"],
                              p["cy"])
        o.start()
        o.run(1e30, steps)
        assert_same(gb.download(k), o.get(), f"instance {k}")
        assert gb.clock(k) == o.clock()
    gb.close()


def test_autotune_restores_state_and_keeps_results():
    pkg, E = ens()
    params = E.ensemble_params(seed=11, n=5)
    print("[gpu] autotune on a 5-instance ensemble", flush=True)
    gb = E.build(params, nlim=12)
    before = [gb.download(k) for k in range(gb.n)]
    clocks = [gb.clock(k) for k in range(gb.n)]
    seg, persist, xcd, ms = gb.autotune(1)
    assert seg in (1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32) and persist in (0, 1, 2, 4, 8, 16, 32)
    assert xcd in (0, 1) and ms > 0
    shape = gb.launch_shape()  # the winner of the re-timed finalists is the shape in effect
    assert shape[:3] == (seg, persist, xcd) and shape[3] in (1, 2, 4)
    for k in range(gb.n):
        assert_same(gb.download(k), before[k], f"restored {k}")
        assert gb.clock(k) == clocks[k]
    gb.run()
    ref = E.build(params, nlim=12)
    ref.run()
    for k in range(gb.n):
        assert_same(gb.download(k), ref.download(k), f"tuned {k}")
    gb.close()
    ref.close()
