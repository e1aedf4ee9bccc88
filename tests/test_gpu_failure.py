"""The per-instance failure word: a state the step produces with a NaN or
infinity, a non-positive density or a pressure below the floor raises its
bits (athmi_get_failure). Once a bit of the instance's stop set is raised the
main-loop step stops on that instance while the others of the batch go on
unchanged. The default stop set leaves the pressure floor out: the reference
floors the pressure and steps on, and so does the run, value for value. The
bad states are ordinary data uploaded from the host."""
import importlib

import numpy as np
import pytest

from helpers import load_pkg

pytestmark = pytest.mark.gpu


def test_failed_instance_stops_and_the_others_go_on():
    pkg = load_pkg()
    E = importlib.import_module(pkg.__name__ + ".ensemble")
    params = E.ensemble_params(seed=9, n=3)
    for p in params:
        p["nx1"], p["nx2"] = 64, 96
        p["x2min"], p["x2max"], p["cy"] = -0.75, 0.75, 0.0
    steps = 12
    print(f"[gpu] failure word: 3 instances, {steps} steps, NaN in #1, E < ME in #2", flush=True)
    gb = E.build(params, nlim=steps)
    gb.set_fail_stop(pkg.FAIL_NONFINITE | pkg.FAIL_DENSITY | pkg.FAIL_PRESSURE)
    gas, b1, b2, b3 = gb.download(1)
    gas[50, 30, 0] = np.nan                 # a NaN density
    gb.upload(1, gas, b1, b2, b3)
    gas, b1, b2, b3 = gb.download(2)
    gas[40:44, 20:24, 4] = -1.0             # energy below the kinetic + magnetic energy
    gb.upload(2, gas, b1, b2, b3)
    assert all(gb.failure(k) == (0, -1) for k in range(3))
    gb.run()  # returns: the failed instances no longer count as running
    f0, n0 = gb.failure(0)
    f1, n1 = gb.failure(1)
    f2, n2 = gb.failure(2)
    print(f"[gpu] failure words {(f0, n0)} {(f1, n1)} {(f2, n2)}; steps "
          f"{[gb.clock(k)[2] for k in range(3)]}", flush=True)
    assert (f0, n0) == (0, -1) and gb.clock(0)[2] == steps
    assert f1 & pkg.FAIL_NONFINITE and n1 == 1 and gb.clock(1)[2] == 1
    assert f2 & pkg.FAIL_PRESSURE and n2 == 1 and gb.clock(2)[2] == 1
    # the healthy instance equals its own plain run
    g1 = E.build([params[0]], nlim=steps)
    g1.run()
    for x, y in zip(gb.download(0), g1.download(0)):
        assert np.array_equal(x, y)
    g1.close()
    # a new state clears the word and the instance steps again
    g2 = E.build([params[1]], nlim=steps)
    gb.upload(1, *g2.download(0))
    gb.set_clock(1, *g2.clock(0))
    assert gb.failure(1) == (0, -1)
    gb.run()
    assert gb.clock(1)[2] == steps and gb.failure(1) == (0, -1)
    g2.close()
    gb.close()


def test_pressure_floor_is_reported_and_the_run_goes_on_as_the_reference():
    """Blast 300x72 (Roe): a cell's pressure reaches the floor at step 9. The
    reference floors it (convert_var.c:49) and steps on; with the default stop
    set the run does the same, equal to the oracle value for value through 16
    steps, and the failure word reports the floor hit and its step."""
    from helpers import oracle_setup
    pkg = load_pkg()
    ov = ["grid/Nx1=300", "grid/Nx2=72", "problem/radius=0.2"]
    print("[gpu] pressure floor: blast 300x72 Roe, 16 steps vs the oracle", flush=True)
    gpu, par, tlim, _ = pkg.setup_problem("blast", ov, variant="exact")
    ora, _, _ = oracle_setup("blast", ov)
    gpu.set_limits(tlim, 16)
    for s in range(16):
        gpu.step()
        ora.step(tlim)
    for x, y in zip(gpu.download(0), ora.get()):
        assert np.array_equal(x, y)
    assert gpu.clock(0) == ora.clock()
    flags, n = gpu.failure(0)
    print(f"[gpu] failure word {(flags, n)}", flush=True)
    assert flags == pkg.FAIL_PRESSURE and 1 <= n <= 16
    # with the floor in the stop set the run stops after the state that hit it
    gpu2, _, _, _ = pkg.setup_problem("blast", ov, variant="exact")
    gpu2.set_fail_stop(pkg.FAIL_NONFINITE | pkg.FAIL_DENSITY | pkg.FAIL_PRESSURE)
    gpu2.set_limits(tlim, 16)
    gpu2.run()
    assert gpu2.failure(0) == (flags, n) and gpu2.clock(0)[2] == n
    gpu.close()
    gpu2.close()
