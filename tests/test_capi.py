"""The C-ABI library: it loads without a GPU, exports every entry point the
headers in include/ declare, and the host-side mirror of the reference's input
surface behaves like par.c. No compute calls (no GPU here)."""
import ctypes
import os
import re

import pytest

from helpers import INPUTS, ROOT, load_pkg


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)) -
                  {"if", "while", "for", "sizeof", "return"})


@pytest.mark.parametrize("header", ["athmi.h", "athena_dropin.h", "athmi_adjoint.h",
                                    "athmi_output.h"])
def test_library_exports_every_declared_entry_point(header):
    pkg = load_pkg()
    lib = ctypes.CDLL(pkg.LIB_PATH)
    names = declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_signatures_cover_the_header():
    pkg = load_pkg()
    names = [n for n in declared("athmi.h")
             if n not in ("athmi_profile_steps", "athmi_step_kernel_name", "athmi_copy_calibration")]
    assert set(names) <= set(pkg.SIGNATURES), set(names) - set(pkg.SIGNATURES)


def test_athinput_parser_and_overrides():
    pkg = load_pkg()
    par = pkg.read_athinput(os.path.join(INPUTS, "athinput.blast"), ["grid/Nx1=64", "problem/angle=30"])
    assert par["grid"]["Nx1"] == "64" and par["problem"]["angle"] == "30"
    assert par["grid"]["ibc_x1"] == "4"
    d = pkg.desc_from_athinput(par)
    assert d["nx1"] == 64 and d["nx2"] == 600 and d["bc"] == (4, 4, 4, 4)
    assert d["gamma"] == 1.6666666666666667 and d["cour_no"] == 0.4
    # par_cmdline (par.c:190-199) rejects overrides of blocks or names not in the file
    with pytest.raises(pkg.AthmiError, match="Block"):
        pkg.read_athinput(os.path.join(INPUTS, "athinput.blast"), ["nosuch/Nx1=3"])
    with pytest.raises(pkg.AthmiError, match="Par"):
        pkg.read_athinput(os.path.join(INPUTS, "athinput.blast"), ["grid/nosuch=3"])


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    pkg = load_pkg()
    monkeypatch.setattr(pkg, "_lib", None)
    monkeypatch.setattr(pkg, "LIB_PATH", str(tmp_path / "libathmi.so"))
    with pytest.raises(pkg.AthmiError, match="missing"):
        pkg.lib()


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU: --gpus above 1 exits with a message, before
    touching any device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
