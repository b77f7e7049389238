"""Build the gfx950 tracer library in-tree: onemirror_amd/libonemirror_amd.so.

    python -m onemirror_amd.build

The library holds the HIP kernels and the C++ host side behind the C ABI of
include/onemirror.h.  Flags that matter for results:
  -ffp-contract=off                      every a*b+c rounds twice, on host and device alike
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division and square root in kernels
  -target-feature -packed-fp32-ops       no packed FP32 instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libonemirror_amd.so")
SOURCES = ["om_scene.cpp", "om_kernels.hip", "om_trace64.hip", "om_batch.cpp", "om_diff.hip", "om_annual.cpp",
           "om_facets.cpp", "om_io.cpp"]
ARCH = os.environ.get("OM_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"),
              shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags() -> list:
    return ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
            "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def build(verbose: bool = False) -> str:
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc()] + flags() + (["-x", "hip"] if src.endswith(".hip") else []) + \
              ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-q" not in sys.argv))
