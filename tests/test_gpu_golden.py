"""The HIP path against the golden fixtures written from the reference's own
CPU sources: the host problem generators (initial states) and whole
main-loop runs, value for value, with the exact (no-contraction) kernels."""
import json
import os

import numpy as np
import pytest

from helpers import load_pkg

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(GOLD, "golden.json")))


@pytest.mark.parametrize("case", META["cases"], ids=lambda c: c["file"][:-4])
def test_gpu_main_loop_matches_golden(case):
    pkg = load_pkg()
    print(f"[gpu] golden {case['file']}", flush=True)
    z = np.load(os.path.join(GOLD, case["file"]))
    g, par, tlim, _ = pkg.setup_problem(case["problem"], case["overrides"], flux=case["flux"],
                                        variant="exact")

    def check(tag):
        a = g.download(0)
        for name, x in zip(("gas", "b1i", "b2i", "b3i"), a):
            ref = z[f"{tag}_{name}"]
            assert np.array_equal(x, ref), f"{tag} {name}: max diff {np.abs(x - ref).max()}"
        assert list(g.clock(0)) == list(z[f"{tag}_clock"]), f"{tag} clock"

    check("s0")
    done = 0
    for c in case["checkpoints"]:
        g.set_limits(tlim, c)
        g.run()
        done = c
        check(f"s{c}")
    g.close()
