"""Build the native libraries in-tree.

    libasr_host.so   C++ host side (node-set builder, terminators), g++ -O3 -fopenmp
    libasr_hip.so    HIP kernels for gfx950 and the chain-engine C ABI, hipcc --offload-arch=gfx950
    fortran/asr_gpu_demo  the Fortran binding (module asr_gpu) and a short caller of the C ABI, amdflang

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


FORTRAN_DIR = os.path.join(PKG, "fortran")
FORTRAN_BUILD = os.path.join(FORTRAN_DIR, "build")
FORTRAN_DEMO = os.path.join(FORTRAN_DIR, "asr_gpu_demo")


def amdflang() -> str | None:
    for c in ("/opt/rocm/bin/amdflang", "/opt/rocm/llvm/bin/flang"):
        if os.path.exists(c):
            return c
    return None


def build_fortran() -> str | None:
    """The Fortran binding (module asr_gpu, asr_amd/fortran/asr_gpu.f90) and its short caller asr_gpu_demo, linked
    against libasr_hip.so (found next to the binary's directory at run time).  Skipped when no Fortran compiler is
    present."""
    fc = amdflang()
    if fc is None:
        print("build_fortran: no amdflang, skipped", flush=True)
        return None
    os.makedirs(FORTRAN_BUILD, exist_ok=True)
    _run([fc, "-O2", "-J", FORTRAN_BUILD, os.path.join(FORTRAN_DIR, "asr_gpu.f90"), os.path.join(FORTRAN_DIR, "asr_gpu_demo.f90"),
          "-L", PKG, "-lasr_hip", "-Wl,-rpath,$ORIGIN/..", "-o", FORTRAN_DEMO + ".tmp"])
    os.replace(FORTRAN_DEMO + ".tmp", FORTRAN_DEMO)
    return FORTRAN_DEMO


def build_all() -> None:
    build_host()
    if all(os.path.exists(s) for s in HIP_SRC):
        build_hip()
        build_fortran()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"host": build_host, "hip": build_hip, "fortran": build_fortran, "all": build_all}[what]()
