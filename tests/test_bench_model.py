"""The bench line's cost model (bench.py): algorithmic bytes and flops per launch of the chain level, the compulsory bytes
of a level (read u_k and u_k-1, write u_k+1, the node tables once), and the roofline figures built from them.

Checked on a small fcc ternary operator: the operator passes move (reads + writes) vectors plus the tables; the
reductions and bookkeeping launches are charged only their per-block partials, so at their measured durations (about
10-25 us on the headline) they stay far below the HBM peak; the launch sequence of the fused level maps to the right
vector counts; a fused level's kernel bound is FP64, not MFMA.
"""
import numpy as np
import pytest

import bench as B
from asr_amd import workloads as W


@pytest.fixture(scope="module")
def op():
    return W.ternary_edc(L=4, levels=2, depth=4).op


def test_operator_pass_bytes_are_vectors_plus_tables(op):
    nb = 54
    vec = op.nn * 9 * nb * 16
    tables = op.nn * (4 * op.nslot + 4 * op.nsmax + 4 + 8)
    by, fl = B.chain_kernel_cost(op, nb, 0, 8, 1, 2)
    assert by == 3 * vec + tables
    by7, fl7 = B.chain_kernel_cost(op, nb, 7, 8, 4, 1)
    assert by7 == 5 * vec + tables and fl7 > fl


def test_reduction_launches_are_charged_their_partials_only(op):
    nb, tile = 54, 8
    vec = op.nn * 9 * nb * 16
    for kind in (2, 4):
        by, fl = B.chain_kernel_cost(op, nb, kind, tile)
        assert by < vec / 50, (kind, by, vec)
        # at 10 us such a launch stays well under the HBM peak
        assert by / 10e-6 / 1e9 < 0.5 * B.HBM_PEAK_GBS


def test_fused_level_launch_sequence(op):
    # first S pass, listed-node terms, a_n, second S pass with the update, b_n
    assert B.level_io([0, 6, 2, 7, 4]) == [B.EDC_IO_FIRST, (1, 1), (1, 1), B.EDC_IO_UPD, (1, 1)]
    # the separate update pass: first and second S pass, a_n, update, b_n
    assert B.level_io([0, 6, 0, 2, 3, 4]) == [B.EDC_IO_FIRST, (1, 1), B.EDC_IO_SECOND, (1, 1), (1, 1), (1, 1)]


def test_compulsory_level_bytes(op):
    nb = 54
    t = 4 * op.nslot + 4 * op.nsmax + 4 + 8
    assert B.compulsory_level_bytes(op, nb) == nb * op.nn * (2 * 9 * 8 * 3) + op.nn * t


def test_roofline_bound_and_fraction():
    r = B.roofline(4.0e9, 4.0e10, 1.0)
    assert r["bound"] == "fp64" and r["unit"] == "TFLOP/s"
    assert abs(r["achieved"] - 40.0) < 1e-9 and abs(r["frac"] - 40.0 / B.FP64_PEAK_TFS) < 1e-3
    r = B.roofline(8.0e9, 1.0e9, 2.0)
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and abs(r["achieved"] - 4000.0) < 1e-6
