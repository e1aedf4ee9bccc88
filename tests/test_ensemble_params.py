"""Host logic of the batched ensemble: the seeded draw is deterministic and
stays inside the ranges the workload states."""
import importlib
import math

from helpers import load_pkg


def test_ensemble_draw():
    pkg = load_pkg()
    E = importlib.import_module(pkg.__name__ + ".ensemble")
    a, b = E.ensemble_params(5, 256), E.ensemble_params(5, 256)
    assert a == b and a != E.ensemble_params(6, 256)
    sizes = set(range(64, 1025, 32))
    for k, p in enumerate(a):
        assert p["nx1"] in sizes and p["nx2"] in sizes
        assert p["flux"] == ("roe" if k % 2 == 0 else "hlle")
        lx, ly = p["x1max"] - p["x1min"], p["x2max"] - p["x2min"]
        assert lx / p["nx1"] == ly / p["nx2"] or math.isclose(lx / p["nx1"], ly / p["nx2"])
        assert abs(p["cx"]) <= 0.3 * lx and abs(p["cy"]) <= 0.3 * ly
        assert 0.05 <= p["radius"] <= 0.15 and 1.0 <= p["prat"] <= 100.0
        assert 0.0 <= p["b0"] <= 1.5 and 0.0 <= p["angle"] <= 180.0
    # the mean zone count is near E[nx]^2 = 544^2
    mean = E.cells(a) / len(a)
    assert 0.8 * 544 ** 2 < mean < 1.25 * 544 ** 2
