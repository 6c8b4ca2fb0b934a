"""Build the package's native libraries in-tree (nav24_amd/_lib/).

  libnav24_ba.so   HIP kernels for gfx950 + host runtime + C ABI (include/*.h)
  libnav24_gen.so  seeded synthetic problem generator (host only)

Run:  python -m nav24_amd.build      (or __graft_entry__.build())
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "_lib")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# gfx950 only; packed-FP32 instructions disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]

HIP_SOURCES = ["ba_kernels.hip", "schur_kernels.hip", "chol_kernels.hip", "pcg_kernels.hip", "nav_ba_host.cpp",
               "cost_kernels.hip",
               "undistort_kernels.hip", "twoview_kernels.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile_hip(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build(verbose: bool = False) -> None:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    gen = os.path.join(LIB, "libnav24_gen.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-shared",
          os.path.join(CSRC, "problem_gen.cpp"), "-o", gen + ".tmp"])
    os.replace(gen + ".tmp", gen)
    with ThreadPoolExecutor(max_workers=len(HIP_SOURCES)) as ex:
        objs = list(ex.map(_compile_hip, HIP_SOURCES))
    ba = os.path.join(LIB, "libnav24_ba.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", ba + ".tmp"])
    os.replace(ba + ".tmp", ba)
    if verbose:
        print("built", gen, ba)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
