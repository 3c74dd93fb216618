"""Build the HIP library lblrtm_amd/liblblx.so for gfx950 (in-tree).

hipcc compiles the one HIP translation unit (kernels + C ABI) and the host
planner; -ffp-contract=off keeps the index and rounding arithmetic the same
as the reference's double-precision build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblblx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LBLX_ARCH", "gfx950")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + cmd[0])


def build(verbose=False):
    bdir = os.path.join(HERE, "_objs")
    os.makedirs(bdir, exist_ok=True)
    plan_o = os.path.join(bdir, "lblx_plan.o")
    api_o = os.path.join(bdir, "lblx_api.o")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-c",
          os.path.join(CSRC, "lblx_plan.cpp"), "-o", plan_o], verbose)
    _run([HIPCC, "-x", "hip", "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
          "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value",
          "-Wno-unused-result", "-c", os.path.join(CSRC, "lblx_api.cpp"), "-o", api_o], verbose)
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", api_o, plan_o,
          "-o", LIB + ".tmp"], verbose)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
