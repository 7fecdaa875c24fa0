"""Builds the native libraries in-tree.

  lib/libssol_hip.so  gfx950 HIP kernels + the ssolh C-ABI   (hipcc)
  lib/libssol.so      the ssol_* C API, host side            (g++, links the above)

Run as a script or call build(). Every call compiles (no staleness check).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "lib")

HIP_SOURCES = [os.path.join(HERE, "csrc", "hip", "ssolh.hip")]
HOST_SOURCES = [os.path.join(HERE, "csrc", "host", f) for f in (
    "ssol_objects.cpp", "ssol_shape.cpp", "ssol_bvh.cpp", "ssol_solve.cpp", "ssol_draw.cpp",
    "ssp_rng.cpp")]

HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
               "-munsafe-fp-atomics", "-fvisibility=hidden",
               # keep packed-FP32 instructions out of the generated code
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               # no machine-level hoisting of loop invariants: it lifts the
               # FP64 polynomial and Threefry constants of the path kernels
               # out of their loops and keeps them live across them, which
               # spills (k_shade 108 -> 60 bytes/lane of scratch, k_generate
               # 92 -> 0, the looping graph-captured forms 388 -> 72)
               "-mllvm", "-disable-machine-licm"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> None:
    os.makedirs(LIB, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    hip_so = os.path.join(LIB, "libssol_hip.so")
    _run([hipcc, *HIPCC_FLAGS, "-I", INCLUDE, "-o", hip_so + ".tmp", *HIP_SOURCES])
    os.replace(hip_so + ".tmp", hip_so)
    host_so = os.path.join(LIB, "libssol.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
          "-DSSOL_SHARED_BUILD", "-Wall", "-Wno-unused-parameter", "-I", INCLUDE,
          "-o", host_so + ".tmp", *HOST_SOURCES, "-L", LIB, "-lssol_hip",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(host_so + ".tmp", host_so)


if __name__ == "__main__":
    build()
    sys.exit(0)
