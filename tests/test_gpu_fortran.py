"""The Fortran caller (module asr_gpu) on the GPU: the operator written in the reference's table layout (augcls_op's
1-based, column-major fields and class_alloy_t's blocks), translated by gpu_op_from_tables and run through the C ABI
from Fortran; the edc and real Lanczos chains against the oracle on the same operator, for the D2h group and for the
reference's D4h group (orbital permutations in pi)."""
import os
import subprocess

import numpy as np
import pytest

from asr_amd import workloads as W

import fortran_layout as FL

pytestmark = pytest.mark.gpu
DEMO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "asr_amd", "fortran", "asr_gpu_demo")


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("group", ["d2h", "d4h"])
def test_fortran_caller_chains_equal_the_oracle(tmp_path, group):
    from oracle.oracle import OracleOp
    assert os.path.exists(DEMO), "asr_amd/fortran/asr_gpu_demo is missing: run build()"
    wl = W.ternary_edc(L=4, levels=2, depth=12, seed=12, group=group)
    l0s = list(range(9)) if group == "d2h" else [0, 2, 4, 6, 8]
    members = [m for m in wl.members if m.l0 in l0s][::3]
    nlev = 12
    fin, fout = str(tmp_path / "op.txt"), str(tmp_path / "out.txt")
    FL.write_reference_layout(fin, wl.op, members, nlev)
    print(f"fortran caller: {group}, {wl.op.nn} nodes, {len(members)} chains", flush=True)
    r = subprocess.run([DEMO, fin, fout], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    mu, a, b, mr, ar, br = FL.read_results(fout, len(members), nlev)
    O = OracleOp(wl.op)
    for j, m in enumerate(members):
        ea, eb, n = O.edc_chain(m, m.z * m.pa + m.pb, m.l0, nlev)
        assert mu[j] == n and rel(a[j, :n], ea[:n]) < 1e-9 and rel(b[j, :n - 1], eb[:n - 1]) < 1e-9, f"edc chain {j}"
        ra, rb, n2 = O.lanczos(m, m.l0, nlev)
        assert mr[j] == n2 and rel(ar[j, :n2], ra[:n2]) < 1e-9 and rel(br[j, :n2], rb[:n2]) < 1e-9, f"real chain {j}"
