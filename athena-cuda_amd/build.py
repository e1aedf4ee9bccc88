"""Build the HIP library of the MHD step for gfx950 (in-tree, no JIT cache).

    python athena-cuda_amd/build.py          # -> athena-cuda_amd/libathmi.so

The step kernels are compiled twice from one source: the "exact" variant with
-ffp-contract=off (one rounding per operation, as the reference CPU build)
and the "fast" variant with FMA contraction. Host code that must reproduce
the reference's host arithmetic (problem generators) is built by g++ with
-ffp-contract=off; the runtime is built by hipcc.
"""
import os
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
OUT = os.path.join(PKG, "libathmi.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "--offload-arch=gfx950"
# packed FP32 instructions stay off on this pool (they have given wrong results)
NO_PK = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-I", INC, "-I", CSRC]


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose_resources=False):
    with tempfile.TemporaryDirectory(prefix="athmi_build_") as tmp:
        objs = []
        for variant, contract in (("exact", "off"), ("fast", "fast")):
            o = os.path.join(tmp, f"step_{variant}.o")
            # the scheduler's register-pressure trackers: the two-wave x1
            # predictor spills less (1.85 against 2.13 ms at 4096^2; the whole
            # step 3% faster)
            # the fast variant also lets division and square root use the
            # reciprocal estimate refined by Newton steps (about an ulp, not
            # correctly rounded): it is held to a tolerance, not to the
            # reference's values
            approx = ["-freciprocal-math", "-fapprox-func"] if variant == "fast" else []
            cmd = [HIPCC, ARCH, *COMMON, *NO_PK, f"-DATHMI_VARIANT={variant}",
                   f"-ffp-contract={contract}", *approx, "-mllvm", "-amdgpu-use-amdgpu-trackers",
                   "-c", os.path.join(CSRC, "step_kernels.hip"), "-o", o]
            if verbose_resources:
                cmd.append("-Rpass-analysis=kernel-resource-usage")
            run(cmd)
            objs.append(o)
        for src in ("athmi_capi.cpp", "calib.hip", "adjoint_host.cpp", "adjoint_capi.cpp",
                    "adjoint_kernels.hip", "output.hip", "pgen.hip", "transfer.hip",
                    "static_step.hip"):
            o = os.path.join(tmp, src.split(".")[0] + ".o")
            # adjoint kernels: the PLM reverse's loops over the seven
            # characteristic fields are fully unrolled (each eigenmatrix entry
            # then folds to a constant case) although their bodies hold reload
            # fences; the scheduler's own register-pressure trackers cut the
            # spills of the Roe and PLM reverses (one wave per SIMD) by 7-35%
            # and their time by about 10% (measured at 512x512)
            extra = (["-mllvm", "-pragma-unroll-threshold=200000",
                      "-mllvm", "-amdgpu-use-amdgpu-trackers"]
                     if src == "adjoint_kernels.hip" else [])
            run([HIPCC, ARCH, *COMMON, *NO_PK, "-ffp-contract=off", *extra, "-c",
                 os.path.join(CSRC, src), "-o", o])
            objs.append(o)
        for src in ("host_problems.cpp", "dropin.cpp"):
            o = os.path.join(tmp, src.replace(".cpp", ".o"))
            run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I", INC, "-I", CSRC,
                 "-c", os.path.join(CSRC, src), "-o", o])
            objs.append(o)
        tmp_out = OUT + ".tmp"
        run([HIPCC, ARCH, "-shared", "-fPIC", "-o", tmp_out, *objs])
        os.replace(tmp_out, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose_resources="-v" in sys.argv)
