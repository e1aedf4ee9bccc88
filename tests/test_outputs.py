"""The output layer on the CPU: the athinput <outputN> parsing and file names
of the host mirror (athena-cuda_amd/output.py), and the oracle's restatement
of the formats (oracle/outputs.py) run with the oracle integrator against the
files the reference's own output layer wrote for the same runs (the SHA-256
manifest of tests/golden/make_outputs.py). No GPU."""
import hashlib
import json
import math
import os
import sys

import numpy as np
import pytest

from helpers import ROOT, OracleGrid, load_pkg

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import outputs as ora_out  # noqa: E402

GOLD = os.path.join(ROOT, "tests", "golden", "outputs")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))
PALETTES = dict(np.load(os.path.join(GOLD, "palettes.npz")))


def test_fname_construct():
    out = load_pkg()
    from athena_cuda_amd import output as O
    assert O.fname_construct("Blast", 4, 7, "d", "ppm") == "Blast.0007.d.ppm"
    assert O.fname_construct("Blast", 4, 12, None, "bin") == "Blast.0012.bin"
    assert O.fname_construct("Blast", 0, 3, None, "bin") == "Blast.bin"
    assert O.fname_construct("cpaw", 4, -1, "V3", "ppm") == "cpaw.V3.ppm"
    assert out is not None


def test_init_output_parses_the_blocks(tmp_path):
    pkg = load_pkg()
    from athena_cuda_amd import output as O
    par = pkg.read_athinput(os.path.join(GOLD, "athinput.blast"))
    outs = O.init_output(par, "blast")
    got = [(o.n, o.out_fmt, o.out, o.id, o.dt, o.sdmin, o.sdmax, o.palette) for o in outs]
    assert got[0] == (1, "ppm", "d", "d", 0.002, 1, 1, "rainbow")
    assert got[1][:3] == (2, "bin", "all")
    assert got[5] == (6, "ppm", "V1", "out6", 0.004, 0, 1, "step32")
    assert [o.out_fmt for o in outs].count("tab") == 2
    assert outs[8].dat_fmt == "%14.6e"
    # a user expression without usr_expr_flag, and an unknown one, are skipped
    # (output.c:244-249); an unknown palette is fatal (output.c:874-880)
    par = pkg.read_athinput(os.path.join(GOLD, "athinput.field_loop"))
    par["output3"]["out"] = "J3"
    assert [o.n for o in O.init_output(par, "field_loop")] == [1, 2, 4]
    par["output3"]["out"] = "nope"
    assert [o.n for o in O.init_output(par, "field_loop")] == [1, 2, 4]
    par["output1"]["palette"] = "grey"
    with pytest.raises(pkg.AthmiError):
        O.init_output(par, "field_loop")


def _sequential_minmax(flat):
    """utils.c:148-162 as written: dmin = dmax = data[0][0], then
    MIN(a,b) = a < b ? a : b, MAX(a,b) = a > b ? a : b over every element."""
    mn = mx = flat[0]
    for x in flat:
        mn = mn if mn < x else x
        mx = mx if mx > x else x
    return mn, mx


@pytest.mark.parametrize("seed", range(6))
def test_minmax2_restatement_is_the_sequential_scan(seed):
    rng = np.random.default_rng(seed)
    a = rng.choice(np.array([-2.0, -0.0, 0.0, 1.0, 3.5], dtype=np.float32), size=(7, 9))
    if seed % 2:
        a.flat[rng.integers(0, a.size, 3)] = np.nan
    if seed == 5:
        a.flat[-1] = np.nan
    mn, mx = ora_out.minmax2(a)
    smn, smx = _sequential_minmax(list(a.ravel()))
    for u, v in ((mn, smn), (mx, smx)):
        assert (math.isnan(u) and math.isnan(v)) or (u == v and math.copysign(1, u) == math.copysign(1, v))


def _oracle_run(problem, outdir):
    """The oracle integrator + the oracle's formats, with main.c's loop and
    output.c's cadence; the <outputN> blocks parsed by the host mirror."""
    pkg = load_pkg()
    from athena_cuda_amd import output as O
    from helpers import oracle_lib
    path = os.path.join(GOLD, f"athinput.{problem}")
    par = pkg.read_athinput(path)
    d = pkg.desc_from_athinput(par)
    g = OracleGrid(d)
    pr, f = par["problem"], pkg._atof
    L = oracle_lib()
    if problem == "blast":
        L.ora_pgen_blast(g.h, f(pr["radius"]), f(pr["pamb"]), f(pr["prat"]), f(pr["b0"]), f(pr["angle"]))
    elif problem == "field_loop":
        L.ora_pgen_field_loop(g.h, f(pr["rad"]), f(pr["amp"]), f(pr["vflow"]), int(f(pr["iprob"])))
    else:
        L.ora_pgen_cpaw2d(g.h, f(pr["angle"]), f(pr["b_par"]), f(pr["b_perp"]), f(pr["pres"]),
                          f(pr["v_par"]))
    g.start()
    tlim, nlim = f(par["time"]["tlim"]), pkg._atoi(par["time"].get("nlim", "-1"))
    outs = O.init_output(par, problem)
    base = par["job"]["problem_id"]
    iv, dv = g.info()
    G = ora_out.Geom(d["nx1"], d["nx2"], dv[0], dv[1], dv[2], dv[3])
    gamma = d["gamma"]
    while True:
        dt, time, nstep = g.clock()
        if not (time < tlim and (nlim < 0 or nstep < nlim)):
            break
        due = []
        for o in outs:
            due.append(time >= o.t)
            if time >= o.t:
                o.t += o.dt
        gas, b1, b2, b3 = g.get()
        for o, fl in zip(outs, due):
            if not fl:
                continue
            if o.out_fmt == "bin":
                ora_out.dump_binary(os.path.join(outdir, ora_out.fname(base, o.num, None, "bin")),
                                    gas, G, time, dt, gamma - 1.0)
            elif o.out_fmt == "tab":
                ora_out.dump_tab(os.path.join(outdir, ora_out.fname(base, o.num, None, "tab")),
                                 gas, b1, b2, b3, G, time, gamma - 1.0, o.dat_fmt)
            else:
                data = ora_out.expression(o.out, gas, b1, b2, G, gamma)
                ora_out.output_ppm(os.path.join(outdir, ora_out.fname(base, o.num, o.id, "ppm")),
                                   data, PALETTES[o.palette], o)
            o.num += 1
        g.step(tlim)
    return g.clock()


def _hashes(d):
    return {n: hashlib.sha256(open(os.path.join(d, n), "rb").read()).hexdigest()
            for n in sorted(os.listdir(d))}


@pytest.mark.parametrize("problem", ["blast", "field_loop", "cpaw2d"])
def test_oracle_outputs_match_the_reference_files(tmp_path, problem):
    dt, time, nstep = _oracle_run(problem, str(tmp_path))
    want = MANIFEST[problem]
    assert nstep == want["clock"]["nstep"] and time == want["clock"]["time"]
    got = _hashes(str(tmp_path))
    assert sorted(got) == sorted(want["files"])
    bad = [n for n in got if got[n] != want["files"][n]["sha256"]]
    assert not bad, bad
