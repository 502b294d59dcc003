"""Build the native libraries in-tree.

    libasr_host.so   C++ host side (node-set builder, terminators), g++ -O3 -fopenmp
    libasr_hip.so    HIP kernels for gfx950 and the chain-engine C ABI, hipcc --offload-arch=gfx950

`python -m asr_amd.build` compiles both every time it is called.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
HOST_SRC = [os.path.join(PKG, "csrc", "host", f) for f in ("augcls_build.cpp", "terminators.cpp")]
HIP_SRC = [os.path.join(PKG, "csrc", "hip", f) for f in ("asr_engine.hip",)] + \
    [os.path.join(PKG, "csrc", "host", "terminators.cpp")]  # cf_species for the adaptive depth test
HOST_LIB = os.path.join(PKG, "libasr_host.so")
HIP_LIB = os.path.join(PKG, "libasr_hip.so")

# Packed FP32 instructions are kept out of every HIP build; -O3 with FP64 contraction left on (fma), no fast-math.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=fast",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_host() -> str:
    _run(["g++", "-O3", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-I", INCLUDE, *HOST_SRC, "-o", HOST_LIB + ".tmp"])
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    return HOST_LIB


def build_hip() -> str:
    _run(["hipcc", *HIP_FLAGS, "-I", INCLUDE, *HIP_SRC, "-o", HIP_LIB + ".tmp"])
    os.replace(HIP_LIB + ".tmp", HIP_LIB)
    return HIP_LIB


def build_all() -> None:
    build_host()
    if all(os.path.exists(s) for s in HIP_SRC):
        build_hip()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"host": build_host, "hip": build_hip, "all": build_all}[what]()
