"""Build libswe_amd.so in place with hipcc for gfx950 (and the host-only mesh generator
libswe_meshgen.so with g++).

    python shallow-water-eric_amd/build_lib.py

-ffp-contract=off keeps every multiply and add separately rounded, as in the reference build.
Packed FP32 instructions are disabled for the whole library.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
OUT = os.path.join(PKG, "libswe_amd.so")
CSRC = os.path.join(PKG, "csrc")
SOURCES = [os.path.join(CSRC, f) for f in ("swe_kernels.hip", "swe_engine.cpp", "swe_capi.cpp", "swe_setup.cpp",
                                                              "swe_setup_device.hip", "swe_errors.hip")]
HEADERS = [os.path.join(REPO, "include", "swe_amd.h")] + [os.path.join(CSRC, f) for f in
                                                          ("swe_kernels.h", "swe_layout.h", "swe_engine.h", "swe_p1.h")]


def hipcc():
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def command(out=OUT, extra=()):
    return [hipcc(), "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(REPO, "include"), "-o", out, *extra, *SOURCES]


# The mesh generator of the benchmark workloads (host C++ only, no GPU code): libswe_meshgen.so.
MESHGEN_OUT = os.path.join(PKG, "libswe_meshgen.so")
MESHGEN_SRC = os.path.join(PKG, "meshgen", "swe_meshgen.cpp")


def meshgen_command(out=MESHGEN_OUT):
    cxx = shutil.which("g++") or "g++"
    return [cxx, "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-o", out, MESHGEN_SRC, "-lpthread"]


def _run(cmd, tmp, out, verbose):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s failed (%d):\n%s\n%s" % (cmd[0], r.returncode, " ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, out)
    if verbose:
        print(" ".join(cmd))


def build(verbose: bool = False) -> str:
    tmp = OUT + ".tmp"
    _run(command(tmp), tmp, OUT, verbose)
    tmp = MESHGEN_OUT + ".tmp"
    _run(meshgen_command(tmp), tmp, MESHGEN_OUT, verbose)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
