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


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")


def build_fortran(verbose=False):
    """The Fortran interface (lblx_c) and the OPDPTH/XMERGE shims, compiled
    with the reference's double-precision flags into lblrtm_amd/fortran/
    _build (objects and module files a maintainer links with the
    reference).  Returns the build directory, or None without flang."""
    if not os.path.exists(FLANG):
        return None
    src = os.path.join(HERE, "fortran")
    out = os.path.join(src, "_build")
    os.makedirs(out, exist_ok=True)
    flags = ["-O1", "-fdefault-integer-8", "-fdefault-real-8", "-fdefault-double-8", "-module-dir", out,
             "-I", out]
    for name in ("lblx_c", "lblx_shims"):
        _run([FLANG, "-c"] + flags + [os.path.join(src, name + ".f90"), "-o", os.path.join(out, name + ".o")],
             verbose)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
