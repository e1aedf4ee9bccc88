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


def test_largest_blast_steps_conserve_and_keep_div_b():
    """The blast at 16384 x 16384: its 81 planes of 16392 x 16400 doubles
    (state, the other ping-pong buffer and the 59 scratch planes) take 174 GB
    of the GPU's 288 GB, and plane offsets reach 2.7e8 of the kernels' 32-bit
    range. Three main-loop steps from the device generator: mass and energy
    conserved to roundoff on the periodic box, div B at roundoff, the clock
    advanced and no failure bit."""
    pkg = load_pkg()
    n = 16384
    print(f"[fullsize] blast {n}x{n}: 3 steps, conservation and div B", flush=True)
    desc = dict(nx1=n, nx2=n, x1min=-0.5, x1max=0.5, x2min=-0.75, x2max=0.75,
                gamma=5.0 / 3.0, cour_no=0.4, bc=(4, 4, 4, 4), flux="roe")
    g = pkg.Grid(desc)
    g.pgen_device(0, [("blast", (0.1, 0.1, 100.0, 1.0, 45.0, 0.0, 0.0))])
    g.set_bvals()
    g.new_dt()
    g.set_limits(1e30, -1)
    dx = 1.0 / n

    def totals():
        gas = g.download(0)[0][4:4 + n, 4:4 + n]
        return np.array([np.sum(gas[..., 0]), np.sum(gas[..., 4])])

    t0 = totals()
    for _ in range(3):
        g.step()
    t1 = totals()
    dt, time, nstep = g.clock(0)
    print(f"[fullsize] {n}^2: nstep {nstep}, time {time:.3e}, mass/energy {t0} -> {t1}",
          flush=True)
    assert nstep == 3 and np.isfinite(dt) and dt > 0.0
    assert g.failure(0) == (0, -1)
    assert np.all(np.abs(t1 - t0) <= 1e-12 * np.abs(t0)), (t0, t1)
    # B is O(1): the roundoff of a few operations per step over dx
    assert g.divb_max(0) * dx <= 3 * 4 * 2.0 ** -52
    g.close()
