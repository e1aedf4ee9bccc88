"""The oracle against the reference's own CPU integrator at full length: the
digests and clocks of tests/golden/fullsize.json (written from oracle/_ref by
tests/golden/make_fullsize.py) for the configurations the oracle runs in
seconds here -- field_loop 128x64 to t = 2 and cpaw2d up to 256^2 for one
period, Roe and HLLE, every 100 main-loop steps. The GPU tests compare the
device against the same digests at every size."""
import json
import os

import pytest

from helpers import oracle_setup
from test_gpu_configs import state_digest

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "fullsize.json")) as _fh:
    FULL = json.load(_fh)
CHEAP = [c for c in FULL["cases"]
         if c["problem"] == "field_loop" or c["name"] in ("cpaw2d_64", "cpaw2d_128", "cpaw2d_256")]


@pytest.mark.parametrize("case", CHEAP, ids=lambda c: f"{c['name']}_{c['flux']}")
def test_oracle_matches_reference_digests_at_full_length(case):
    g, tlim, _ = oracle_setup(case["problem"], case["overrides"], flux=case["flux"])
    done = 0
    for cp in case["checkpoints"]:
        if cp["nstep"] > done:
            g.run(tlim, cp["nstep"] - done)
            done = cp["nstep"]
        dt, time, nstep = g.clock()
        assert (nstep, dt.hex(), time.hex()) == (cp["nstep"], cp["dt"], cp["time"])
        assert state_digest(*g.get()) == cp["sha256"], f"step {cp['nstep']}"
    assert g.clock()[1] == tlim
