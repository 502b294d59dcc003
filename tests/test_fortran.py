"""The Fortran binding (asr_amd/fortran/asr_gpu.f90, module asr_gpu) and its caller, built by amdflang and linked
against libasr_hip.so (no GPU needed here):

* the caller runs, finds no GPU, and a call the engine refuses returns its status and message through the binding;
* the module that takes the reference's own types (asr_gpu_augcls: augcls_op, class_alloy_t, edn_t, cf_species) compiles
  against the reference's modules of the chain path (compile-only, oracle/_ref/fmod, where the reference's sources are
  present).
The GPU run of the same caller on tables in the reference's layout is tests/test_gpu_fortran.py."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "asr_amd", "fortran", "asr_gpu_demo")


@pytest.mark.skipif(not os.path.exists(DEMO), reason="the Fortran caller is built by build() where amdflang exists")
def test_fortran_caller_links_and_reports_the_engine_error():
    env = dict(os.environ, HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", ""))
    r = subprocess.run([DEMO], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0].startswith("devices ")
    assert "status 1 message asr_gpu_op_create: dimensions out of range" in r.stdout


@pytest.mark.skipif(not (os.path.isdir("/root/reference/src") and os.path.exists("/opt/rocm/bin/amdflang")),
                    reason="needs the reference's Fortran sources and amdflang (build container only)")
def test_fortran_augcls_binding_compiles_against_the_reference_modules():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(os.path.join(ROOT, "oracle", "_ref", "fmod", "asr_gpu_augcls.o"))
