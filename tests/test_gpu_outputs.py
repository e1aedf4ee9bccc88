"""The output layer on the GPU, through the C-ABI (include/athmi_output.h)
and the host mirror of output.c / main.c (athena-cuda_amd/output.py):

* whole runs of the small problems of tests/golden/outputs/ write exactly
  the files, byte for byte, that the reference's own output layer wrote
  (SHA-256 manifest from tests/golden/make_outputs.py);
* at larger sizes, every expression through every palette, with automatic,
  fixed and inverted scales, plus the binary and table dumps, equal the
  oracle's restatement (oracle/outputs.py) applied to the same device state;
* the min/max scan keeps the reference's NaN and signed-zero behaviour.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from helpers import ROOT, load_pkg

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import outputs as ora_out  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = os.path.join(ROOT, "tests", "golden", "outputs")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))
PALETTES = dict(np.load(os.path.join(GOLD, "palettes.npz")))
EXPRS = ("d", "M1", "M2", "M3", "E", "B1c", "B2c", "B3c", "ME", "V1", "V2", "V3", "P", "cs2", "S")


def _hashes(d):
    return {n: hashlib.sha256(open(os.path.join(d, n), "rb").read()).hexdigest()
            for n in sorted(os.listdir(d))}


@pytest.mark.parametrize("problem", ["blast", "field_loop", "cpaw2d"])
def test_main_loop_outputs_equal_the_reference_files(tmp_path, problem):
    pkg = load_pkg()
    from athena_cuda_amd import output as O
    print(f"[outputs] {problem}", flush=True)
    g, par, tlim, nlim = pkg.setup_problem(problem, athinput=os.path.join(GOLD, f"athinput.{problem}"))
    outs = O.init_output(par, problem)
    steps, files = O.run_main_loop(g, outs, tlim, nlim, par["job"]["problem_id"], str(tmp_path))
    want = MANIFEST[problem]
    dt, time, nstep = g.clock()
    g.close()
    assert nstep == want["clock"]["nstep"] and time == want["clock"]["time"]
    assert steps == want["clock"]["steps"]
    got = _hashes(str(tmp_path))
    assert sorted(got) == sorted(want["files"])
    bad = [n for n in got if got[n] != want["files"][n]["sha256"]]
    assert not bad, bad


class _Out:
    def __init__(self, num, sdmin=0, dmin=0.0, sdmax=0, dmax=0.0):
        self.num, self.sdmin, self.dmin, self.sdmax, self.dmax = num, sdmin, dmin, sdmax, dmax
        self.gmin = self.gmax = 0.0


def _device_and_oracle_ppm(g, O, tmp_path, expr, palette, out, G, gamma):
    spec = O._PpmSpec(expr.encode(), palette.encode(), out.sdmin, out.sdmax, out.dmin, out.dmax)
    import ctypes
    gmin, gmax = ctypes.c_double(out.gmin), ctypes.c_double(out.gmax)
    p_dev = str(tmp_path / f"dev.{expr}.{palette}.ppm")
    L = O._bind()
    assert L.athmi_output_ppm(g._h, 0, p_dev.encode(), ctypes.byref(spec), out.num,
                              ctypes.byref(gmin), ctypes.byref(gmax)) == 0
    gas, b1, b2, b3 = g.download()
    data = ora_out.expression(expr, gas, b1, b2, G, gamma)
    p_ora = str(tmp_path / f"ora.{expr}.{palette}.ppm")
    ora_out.output_ppm(p_ora, data, PALETTES[palette], out)
    assert repr((gmin.value, gmax.value)) == repr((out.gmin, out.gmax))  # NaN and -0 aware
    return open(p_dev, "rb").read(), open(p_ora, "rb").read()


def test_formats_equal_the_oracle_at_size(tmp_path):
    pkg = load_pkg()
    from athena_cuda_amd import output as O
    print("[outputs] blast 256x384 formats", flush=True)
    g, par, tlim, nlim = pkg.setup_problem("blast", ["grid/Nx1=256", "grid/Nx2=384"])
    g.run(40)
    d = g.descs[0]
    gamma = d["gamma"]
    G = ora_out.Geom(d["nx1"], d["nx2"], d["x1min"], d["x2min"],
                     (d["x1max"] - d["x1min"]) / d["nx1"], (d["x2max"] - d["x2min"]) / d["nx2"])
    dt, time, nstep = g.clock()
    gas, b1, b2, b3 = g.download()
    # binary and table dumps
    o = O.Output(2, 0.1, 0.0, 3, "all", "out2", "bin")
    path = O.write_output(g, o, "Blast", str(tmp_path))
    ora_out.dump_binary(str(tmp_path / "ora.bin"), gas, G, time, dt, gamma - 1.0)
    assert open(path, "rb").read() == open(tmp_path / "ora.bin", "rb").read()
    for fmt in (None, "%14.6e", "%g"):
        o = O.Output(3, 0.1, 0.0, 0, "all", "out3", "tab")
        o.dat_fmt = fmt
        path = O.write_output(g, o, "Blast", str(tmp_path))
        ora_out.dump_tab(str(tmp_path / "ora.tab"), gas, b1, b2, b3, G, time, gamma - 1.0, fmt)
        assert open(path).read() == open(tmp_path / "ora.tab").read(), fmt
    # every expression through every palette, with several scales
    palettes = sorted(PALETTES)
    for k, expr in enumerate(EXPRS):
        pal = palettes[k % len(palettes)]
        for out in (_Out(0), _Out(1, sdmin=1, dmin=-0.3), _Out(2, sdmax=1, dmax=0.7),
                    _Out(0, 1, 2.0, 1, -1.0)):
            a, b = _device_and_oracle_ppm(g, O, tmp_path, expr, pal, out, G, gamma)
            assert a == b, (expr, pal, out.num, out.sdmin, out.sdmax)
    g.close()


def test_field_loop_user_expressions_at_size(tmp_path):
    pkg = load_pkg()
    from athena_cuda_amd import output as O
    print("[outputs] field_loop 256x128 J3 / Bp2", flush=True)
    g, par, tlim, nlim = pkg.setup_problem("field_loop", ["grid/Nx1=256", "grid/Nx2=128"])
    g.run(20)
    d = g.descs[0]
    G = ora_out.Geom(d["nx1"], d["nx2"], d["x1min"], d["x2min"],
                     (d["x1max"] - d["x1min"]) / d["nx1"], (d["x2max"] - d["x2min"]) / d["nx2"])
    for expr in ("J3", "Bp2"):
        for pal in ("rainbow", "jh_colors"):
            for out in (_Out(0), _Out(0, 1, -0.04, 1, 0.08)):
                a, b = _device_and_oracle_ppm(g, O, tmp_path, expr, pal, out, G, d["gamma"])
                assert a == b, (expr, pal)
    g.close()


def test_min_max_scan_keeps_nan_restart_and_signed_zero_ties(tmp_path):
    """minmax2 (utils.c:148-162): a NaN restarts the scan; among equal values
    the later one stays, so -0 and +0 are told apart in the header."""
    pkg = load_pkg()
    from athena_cuda_amd import output as O
    print("[outputs] min/max scan edge cases", flush=True)
    nx1, nx2 = 96, 70
    g = pkg.Grid(dict(nx1=nx1, nx2=nx2, x1min=0.0, x1max=1.0, x2min=0.0, x2max=1.0,
                      gamma=5.0 / 3.0, cour_no=0.4))
    G = ora_out.Geom(nx1, nx2, 0.0, 0.0, 1.0 / nx1, 1.0 / nx2)
    rng = np.random.default_rng(3)
    sx, sy = nx1 + 8, nx2 + 8
    for case in range(4):
        gas = np.zeros((sy, sx, 8))
        gas[..., 0] = rng.choice([0.0, -0.0, 0.5, 2.0], size=(sy, sx))
        gas[..., 1] = rng.choice([-0.0, 0.0, 1.0], size=(sy, sx))
        if case >= 1:
            gas[4 + rng.integers(0, nx2, 5), 4 + rng.integers(0, nx1, 5), 0] = np.nan
        if case == 2:
            gas[4 + nx2 - 1, 4 + nx1 - 1, 0] = np.nan  # the scan ends on a NaN
        if case == 3:
            gas[4 + nx2 - 1, 4 + nx1 - 1, 1] = -0.0
        z = np.zeros((sy, sx))
        g.upload(0, gas, z, z, z)
        for expr in ("d", "M1"):
            a, b = _device_and_oracle_ppm(g, O, tmp_path, expr, "heat", _Out(0), G, 5.0 / 3.0)
            assert a == b, (case, expr)
    g.close()
