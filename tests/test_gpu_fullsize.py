"""The benchmark's own configuration at its real size: the MHD blast at
4096 x 4096 (athinput.blast scaled, Roe + PLM, CTU+CT, periodic), five
main-loop steps on the GPU against the oracle (OpenMP on the host), every
conserved and face field and the clock bit for bit; then twenty more steps
with the size-independent properties of the scheme (mass and energy
conserved to roundoff on the periodic box, div B at roundoff)."""
import numpy as np
import pytest

from helpers import load_pkg, oracle_setup

pytestmark = pytest.mark.gpu
OV = ["grid/Nx1=4096", "grid/Nx2=4096"]


def test_bench_configuration_matches_the_oracle_bit_for_bit():
    pkg = load_pkg()
    print("[fullsize] blast 4096x4096, 5 steps, GPU vs oracle", flush=True)
    g, par, tlim, _ = pkg.setup_problem("blast", OV)
    g.set_limits(tlim, 5)
    g.run()
    gpu = g.download(0)
    clock = g.clock()
    g.close()
    ora, tlim_o, _ = oracle_setup("blast", OV)
    ora.run(tlim_o, 5)
    ref = ora.get()
    for name, a, b in zip(("gas", "b1i", "b2i", "b3i"), gpu, ref):
        assert np.array_equal(a, b), f"{name}: max |diff| {np.nanmax(np.abs(a - b))}"
    assert tuple(clock) == tuple(ora.clock())


def test_bench_configuration_conserves_and_keeps_div_b():
    pkg = load_pkg()
    print("[fullsize] blast 4096x4096, 20 steps, conservation and div B", flush=True)
    g, par, tlim, _ = pkg.setup_problem("blast", OV)
    g.set_limits(tlim, -1)
    d = g.descs[0]
    n1, n2 = d["nx1"], d["nx2"]
    dx = min((d["x1max"] - d["x1min"]) / n1, (d["x2max"] - d["x2min"]) / n2)

    def totals():
        gas = g.download(0)[0][4:4 + n2, 4:4 + n1]
        return np.array([np.sum(gas[..., 0]), np.sum(gas[..., 4])])

    t0 = totals()
    for s in range(20):
        g.step()
        if s % 5 == 4:
            gas, b1, b2, b3 = g.download(0)
            scale = max(np.abs(b1).max(), np.abs(b2).max()) / dx
            assert g.divb_max(0) / scale <= 1e-14, s
    t1 = totals()
    assert np.all(np.abs(t1 - t0) <= 1e-12 * np.abs(t0)), (t0, t1)
    g.close()
