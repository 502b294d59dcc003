"""GPU parity of route ed3 (asr_gpu_ed3_blocks; ed3_blocks / ed3_blocks_blk, src/asr_ed3.f90:247-340) through the C ABI.

Bar: the two chains' a_n, b_n and projections c_n to 1e-9 relative against the oracle on the same seeded inputs, the central
blocks bk to 1e-8 relative; on the ring the blocks equal the exact configuration averages of G_11, (hG)_11, (hGh)_11 of
K = E O - H (1e-6, chains run until the Krylov space is exhausted, no tail).
"""
import numpy as np
import pytest

from asr_amd import ed3 as E3
from asr_amd import workloads as W

pytestmark = pytest.mark.gpu


def _engine():
    from asr_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    return engine


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300))


@pytest.fixture(scope="module")
def ed3_small():
    return W.ternary_ed3(L=4, levels=2, depth=14, seed=21)


def _oracle_blocks(O, wl, j, nlev, tail="hull"):
    m = wl.members[j]
    et, uu, qt = E3.ed3_coefs(wl.info["ed"][j], m.z)
    return O.ed3_blocks(m, et, uu, qt, wl.info["oc"][j], m.z, m.l0, nlev, project=True, tail=tail)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_ed3_blocks_match_oracle(ed3_small, mode):
    """Batched (0), plain (1) and graph (2) runs of 18 ed3 chains (both spins, every central species, several start
    orbitals) against the oracle's ed3_blocks: coefficients, projections, stopping levels and the blocks."""
    eng = _engine()
    from oracle.oracle import OracleOp
    wl = ed3_small
    sel = list(range(0, len(wl.members), 3))
    members = [wl.members[j] for j in sel]
    nlev = 14
    G = eng.GpuOperator(wl.op)
    G.set_depth_tol(0.0, "hull")
    print(f"ed3 mode {mode}: {len(sel)} chains, {wl.op.nn} nodes", flush=True)
    bk, a, b, c, mu, _ = G.ed3_blocks(members, wl.info["cf"][sel], wl.info["oc"][sel], nlev, mode=mode)
    O = OracleOp(wl.op)
    for q, j in enumerate(sel):
        bko, chains, _ = _oracle_blocks(O, wl, j, nlev)
        for sd in range(2):
            ao, bo, co, mo, closed = chains[sd]
            assert not closed
            assert mu[q, sd] == mo
            assert rel(a[q, sd, :mo], ao[:mo]) < 1e-9, (j, sd)
            assert rel(b[q, sd, :mo - 1], bo[:mo - 1]) < 1e-9, (j, sd)
            assert rel(c[q, sd, :mo], co[:mo]) < 1e-9, (j, sd)
        assert rel(bk[q], bko) < 1e-8, (j, bk[q], bko)


def test_ed3_graph_replays_and_modes_agree(ed3_small):
    """Two graph runs are bitwise equal and equal the host-issued batched run (same member tile)."""
    eng = _engine()
    wl = ed3_small
    sel = list(range(0, len(wl.members), 2))
    members = [wl.members[j] for j in sel]
    G = eng.GpuOperator(wl.op)
    r0 = G.ed3_blocks(members, wl.info["cf"][sel], wl.info["oc"][sel], 14, mode=0)
    r1 = G.ed3_blocks(members, wl.info["cf"][sel], wl.info["oc"][sel], 14, mode=2)
    r2 = G.ed3_blocks(members, wl.info["cf"][sel], wl.info["oc"][sel], 14, mode=2)
    for k in range(5):
        assert np.array_equal(r1[k], r2[k])
        assert np.array_equal(r0[k], r1[k])


def test_ed3_ring_exact_average_on_gpu():
    """The ring of test/test_asr_ed3.f90 through the GPU: the blocks of the exhausted chains equal the configuration
    averages of G_11, (hG)_11 and (hGh)_11 of K = E O - H computed by brute force; the two cross terms agree."""
    eng = _engine()
    from exact import build_ring_operator, exact_average_asa, ring_model
    from test_oracle_exact import ring_ed3_params, rotated
    model, _ = ring_model()
    model = rotated(model, 2)
    op = build_ring_operator(model, frozen_start=True)
    pp = ring_ed3_params()
    orb = {k: E3.per_orbital(pp[k]) for k in ("c", "enu", "o", "p")}
    orb["dl"] = E3.per_orbital(pp["sqrdel"]) ** 2
    members, cf, oc, cases = [], [], [], []
    for cen in (0, 3):
        ed = E3.ed3_setup(pp, cen)
        for l0 in (0, 2, 6):
            z = complex(-0.2 + 0.15 * l0 / 6, 0.3)
            members.append(E3.ed3_member(ed, l0, z))
            cf.append(E3.ed3_linear(ed))
            oc.append(E3.ed3_oc(ed, l0))
            cases.append((cen, l0, z))
    G = eng.GpuOperator(op)
    G.set_depth_tol(0.0, "none")
    print(f"ed3 ring: {len(members)} chains, {op.nn} nodes, 1500 levels", flush=True)
    bk, a, b, c, mu, _ = G.ed3_blocks(members, np.array(cf), np.array(oc), 1500, mode=2)
    for q, (cen, l0, z) in enumerate(cases):
        exact = np.array(exact_average_asa(model, orb, z, l0, cen))
        assert np.max(np.abs(bk[q, :3] - exact)) < 1e-6, (cen, l0, bk[q, :3], exact)
        assert abs(bk[q, 3]) < 1e-6


def test_ed3_adaptive_depth():
    """Adaptive depth of route ed3 (chain, src/asr_ed3.f90:177-184): at levels 8, 12, ... the chain terminated with the
    hull tail leaves when |G_k - G_{k-4}| <= tol |G_k| and |f_k - f_{k-4}| <= tol max(|f_k|, |t| |G_k|).  The stopping
    levels are those of the rule on the chains' own coefficients; the coefficients up to there are those of the full run;
    the blocks are the combination of the truncated chains."""
    eng = _engine()
    from asr_amd import host
    from oracle.oracle import OracleOp
    wl = W.ternary_ed3(L=4, levels=2, depth=24, seed=22)
    sel = list(range(0, len(wl.members), 3))
    members = [wl.members[j] for j in sel]
    cf, oc = wl.info["cf"][sel], wl.info["oc"][sel]
    nlev = 24
    G = eng.GpuOperator(wl.op)
    G.set_depth_tol(0.0, "hull")
    bk0, a0, b0, c0, mu0, _ = G.ed3_blocks(members, cf, oc, nlev)
    assert np.all(mu0 == nlev)
    O = OracleOp(wl.op)
    norms = []
    for q, m in enumerate(members):
        _, _, nv, na = O.ed3_start(m, m.l0, oc[q], project=True)
        norms.append(((na, nv), (nv, na)))  # (start norm, projection norm) per side

    def gf(a, b, c, k1, z):
        return host.cf_species(a, b, c, k1, z, "hull")

    ch = np.array([abs(gf(a0[q, sd], b0[q, sd], c0[q, sd], 12, m.z)[0] - gf(a0[q, sd], b0[q, sd], c0[q, sd], 8, m.z)[0])
                   / abs(gf(a0[q, sd], b0[q, sd], c0[q, sd], 12, m.z)[0]) for q, m in enumerate(members) for sd in range(2)])
    u = np.unique(np.round(ch, 12))
    mid = len(u) // 2
    lo, hi = u[mid - 1], u[mid]
    assert hi > lo * (1 + 1e-6)
    tol = float(np.sqrt(lo * hi))
    G.set_depth_tol(tol, "hull")
    bk, a, b, c, mu, _ = G.ed3_blocks(members, cf, oc, nlev)
    G.set_depth_tol(0.0, "hull")
    for q, m in enumerate(members):
        res = []
        for sd in range(2):
            tn = norms[q][sd][1]
            stop, gp, fp = nlev, None, None
            for k1 in range(8, nlev, 4):
                g, f = gf(a[q, sd], b[q, sd], c[q, sd], k1, m.z)
                if k1 > 8 and abs(g - gp) <= tol * abs(g) and abs(f - fp) <= tol * max(abs(f), tn * abs(g)):
                    stop = k1
                    break
                gp, fp = g, f
            assert mu[q, sd] == stop, (q, sd, mu[q, sd], stop)
            k = int(mu[q, sd])
            assert np.array_equal(a[q, sd, :k], a0[q, sd, :k])
            assert np.array_equal(b[q, sd, :k - 1], b0[q, sd, :k - 1])
            assert np.array_equal(c[q, sd, :k], c0[q, sd, :k])
            assert np.all(a[q, sd, k:] == 0.0)
            g, f = gf(a[q, sd], b[q, sd], c[q, sd], k, m.z)
            nr = norms[q][sd][0]
            res.append((nr ** 2 * g, nr * f))
        (q0, q1a), (q2, q1b) = res
        fm = 0.5 * (q1a + q1b)
        want = np.array([q0 - 2 * oc[q] * fm + oc[q] ** 2 * q2, fm - oc[q] * q2, q2, q1a - q1b])
        assert rel(bk[q], want) < 1e-10
    assert np.any(mu == 12) and np.any(mu > 12)
