"""Writes cr_eg_chains.json: the two energy-dependent Cr e_g chains (Re a, Im a, Re b, Im b per level, 24 levels) and
their contour energies that the reference's terminator test holds (test/test_terminators.f90:25-82, arrays cr1, cr2 and
the parameters zc1, zc2).  Reads the test file as text; nothing of the reference is imported or run.

    python tests/golden/make_terminator_chains.py /root/reference/test/test_terminators.f90
"""
import json
import os
import re
import sys


def _numbers(block: str):
    return [float(t.replace("_dp", "").replace("d", "e")) for t in re.findall(r"[-+]?\d\.\d+e[-+]\d+_dp", block)]


def main(path: str):
    src = open(path).read()
    out = {}
    for name in ("cr1", "cr2"):
        m = re.search(name + r"\(4, 24\) = reshape\(\[(.*?)\], \[4, 24\]\)", src, re.S)
        v = _numbers(m.group(1))
        assert len(v) == 96, (name, len(v))
        out[name] = [v[4 * k:4 * k + 4] for k in range(24)]
    for name in ("zc1", "zc2"):
        m = re.search(name + r" = cmplx\(([^,]+), ([^,]+), dp\)", src)
        out[name] = [float(m.group(1).replace("_dp", "")), float(m.group(2).replace("_dp", ""))]
    out["source"] = "test/test_terminators.f90:25-82 (cr1, cr2: columns Re a, Im a, Re b, Im b; zc1, zc2)"
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cr_eg_chains.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test/test_terminators.f90")
