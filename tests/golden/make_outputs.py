"""Writes the golden vectors of the output formats: runs the reference's own
main loop with its output layer (init_output / data_output and dump_binary,
dump_tab, output_ppm; compiled from the reference sources into oracle/_ref)
on the small problems of tests/golden/outputs/athinput.*, and records the
SHA-256 of every file it writes, with the final clock, in
tests/golden/outputs/manifest.json. Also stores the seven palettes as the
reference's getRGB returns them (tests/golden/outputs/palettes.npz).

    make -C oracle ref && python tests/golden/make_outputs.py
"""
import ctypes
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "libathref_roe.so")
CASES = {"blast": 0, "field_loop": 1, "cpaw2d": 2}
PALETTES = ("rainbow", "jh_colors", "idl1", "idl2", "step8", "step32", "heat")

RUN = r"""
import ctypes, json, os, sys
L = ctypes.CDLL(sys.argv[1])
L.ref_open(sys.argv[2].encode(), 0, (ctypes.c_char_p * 1)())
L.ref_set_outfilename(sys.argv[3].encode())
L.ref_setup(int(sys.argv[4]))
L.ref_init_output()
n = L.ref_run_with_output(ctypes.c_double(float(sys.argv[5])), int(sys.argv[6]))
iv = (ctypes.c_int * 9)(); dv = (ctypes.c_double * 6)(); L.ref_info(iv, dv)
json.dump({"steps": n, "nstep": iv[6], "dt": dv[4], "time": dv[5]}, open(sys.argv[7], "w"))
"""


def par_value(path, block, name):
    cur = None
    for raw in open(path):
        line = raw.split("#", 1)[0].strip()
        if line.startswith("<"):
            cur = line[1:line.index(">")]
        elif cur == block and "=" in line and line.split("=", 1)[0].strip() == name:
            return line.split("=", 1)[1].strip()
    raise KeyError(name)


def run_case(problem, outdir):
    inp = os.path.join(HERE, "outputs", f"athinput.{problem}")
    clock = os.path.join(outdir, "clock.json")
    subprocess.run([sys.executable, "-c", RUN, REF, inp, par_value(inp, "job", "problem_id"),
                    str(CASES[problem]), par_value(inp, "time", "tlim"),
                    par_value(inp, "time", "nlim"), clock],
                   check=True, cwd=outdir, stdout=subprocess.DEVNULL)
    info = json.load(open(clock))
    os.remove(clock)
    files = {}
    for name in sorted(os.listdir(outdir)):
        data = open(os.path.join(outdir, name), "rb").read()
        files[name] = {"sha256": hashlib.sha256(data).hexdigest(), "bytes": len(data)}
    return info, files


def main():
    manifest = {}
    for problem in CASES:
        with tempfile.TemporaryDirectory() as d:
            info, files = run_case(problem, d)
        manifest[problem] = {"clock": info, "files": files}
        print(problem, info, len(files), "files")
    json.dump(manifest, open(os.path.join(HERE, "outputs", "manifest.json"), "w"), indent=1,
              sort_keys=True)
    L = ctypes.CDLL(REF)
    pal = {}
    for name in PALETTES:
        a = np.zeros(768, dtype=np.float32)
        L.ref_get_palette(name.encode(), a.ctypes.data_as(ctypes.c_void_p))
        pal[name] = a
    np.savez(os.path.join(HERE, "outputs", "palettes.npz"), **pal)


if __name__ == "__main__":
    main()
