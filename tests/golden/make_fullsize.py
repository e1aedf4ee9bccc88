"""Write tests/golden/fullsize.json: digests of the reference's own CPU
integrator (oracle/_ref, compiled from /root/reference/src by oracle/Makefile;
this container only) run at the stated sizes and lengths of the workloads:

  blast 400x600 on [-0.5,0.5]x[-0.75,0.75] to t = 0.2     (Roe, HLLE)
  field_loop 128x64 on [-1,1]x[-0.5,0.5] to t = 2          (Roe, HLLE)
  cpaw2d N x N, N = 64 .. 1024, to t = 1 (one wave period) (Roe, HLLE)

The full states are too large to commit, and the GPU's exact variant equals
the reference value for value, so each checkpoint stores the SHA-256 of the
padded state (Gas AoS [sy][sx][8], then B1i, B2i, B3i [sy][sx], little-endian
float64 with -0.0 folded into +0.0, since value equality does not see the sign
of zero) together with the clock (dt, time as float.hex, nstep) and the
per-field sums over the padded grid (diagnostics only: they name the field
that moved when a digest differs).

Run: python tests/golden/make_fullsize.py [-j N]   (needs oracle/_ref built)
"""
import ctypes
import hashlib
import json
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
INPUTS = os.path.join(ROOT, "athena-cuda_amd", "inputs")
D, I, P = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
PROBLEM_ID = {"blast": 0, "field_loop": 1, "cpaw2d": 2}

# (name, problem, flux, overrides, checkpoint spacing in main-loop steps)
CASES = [("blast_400x600", "blast", f, [], 100) for f in ("roe", "hlle")]
CASES += [("field_loop_128x64", "field_loop", f, [], 100) for f in ("roe", "hlle")]
CASES += [(f"cpaw2d_{n}", "cpaw2d", f, [f"grid/Nx1={n}", f"grid/Nx2={n}"], 100)
          for n in (64, 128, 256, 512) for f in ("roe", "hlle")]
# one period at 1024^2 (about 2000 steps; half an hour per flux on one core)
CASES += [("cpaw2d_1024", "cpaw2d", f, ["grid/Nx1=1024", "grid/Nx2=1024"], 1024)
          for f in ("roe", "hlle")]


def state_digest(gas, b1i, b2i, b3i):
    """SHA-256 of the padded state, -0.0 folded into +0.0."""
    h = hashlib.sha256()
    for a in (gas, b1i, b2i, b3i):
        h.update(np.ascontiguousarray(np.asarray(a, dtype="<f8") + 0.0).tobytes())
    return h.hexdigest()


def field_sums(gas, b1i, b2i, b3i):
    return [float(np.sum(gas[..., f])) for f in range(8)] + [float(np.sum(b)) for b in
                                                              (b1i, b2i, b3i)]


def one_case(k):
    """Run in a fresh process: the reference keeps its boundary functions in
    file-scope statics, so each configuration needs its own library instance."""
    name, problem, flux, overrides, every = CASES[k]
    L = ctypes.CDLL(os.path.join(REF, f"libathref_{flux}.so"))
    L.ref_run.argtypes = [D, I]
    ov = [o.encode() for o in overrides]
    arr = (ctypes.c_char_p * max(1, len(ov)))(*ov)
    L.ref_open(os.path.join(INPUTS, f"athinput.{problem}").encode(), len(ov), arr)
    L.ref_setup(PROBLEM_ID[problem])
    iv = (I * 9)()
    dv = (D * 6)()
    L.ref_info(iv, dv)
    sx, sy = iv[0] + 8, iv[1] + 8
    tlim = {"blast": 0.2, "field_loop": 2.0, "cpaw2d": 1.0}[problem]
    checkpoints = []

    def snap():
        U = np.zeros((sy, sx, 8))
        B = [np.zeros((sy, sx)) for _ in range(3)]
        L.ref_get(U.ctypes.data_as(P), *[b.ctypes.data_as(P) for b in B])
        L.ref_info(iv, dv)
        checkpoints.append({"nstep": int(iv[6]), "dt": float(dv[4]).hex(),
                            "time": float(dv[5]).hex(), "sha256": state_digest(U, *B),
                            "sums": field_sums(U, *B)})

    snap()
    while True:
        n = L.ref_run(D(tlim), every)
        if n == 0:
            break
        snap()
    return {"name": name, "problem": problem, "flux": flux, "overrides": overrides,
            "tlim": tlim, "shape": [sy, sx], "steps": checkpoints[-1]["nstep"],
            "checkpoints": checkpoints}


def main():
    jobs = 4
    if "-j" in sys.argv:
        jobs = int(sys.argv[sys.argv.index("-j") + 1])

    def run(k):
        r = subprocess.run([sys.executable, __file__, "--case", str(k)], capture_output=True,
                           text=True, check=True)
        line = [x for x in r.stdout.splitlines() if x.startswith("RESULT ")][-1]
        res = json.loads(line[len("RESULT "):])
        print(f"{res['name']} {res['flux']}: {res['steps']} steps", flush=True)
        return res

    # the longest cases first, so that they overlap the short ones
    order = sorted(range(len(CASES)), key=lambda k: "1024" not in CASES[k][0])
    with ThreadPoolExecutor(jobs) as ex:
        done = dict(zip(order, ex.map(run, order)))
    cases = [done[k] for k in range(len(CASES))]
    meta = {"source": "reference CPU sources compiled by oracle/Makefile (oracle/_ref), "
                      "main loop of src/main.c:221-318 (ONLY_HOST)",
            "digest": "sha256 over Gas[sy][sx][8], B1i, B2i, B3i [sy][sx] as <f8, -0.0 -> +0.0",
            "cases": cases}
    with open(os.path.join(HERE, "fullsize.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    print("wrote", len(cases), "cases to fullsize.json")


if __name__ == "__main__":
    if len(sys.argv) > 2 and sys.argv[1] == "--case":
        print("RESULT " + json.dumps(one_case(int(sys.argv[2]))), flush=True)
        sys.exit(0)
    main()
