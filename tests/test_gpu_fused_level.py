"""The fused level of the edc chains (asr_gpu_op_set_level_form 0, the default on the member-tile pass) against the
oracle's edc_green (src/asr_edc.f90:95-155, whose loop applies the operator, takes a_n = sum(u w), then updates and
projects) and against the separate update pass (form 1).

a_n of the fused level comes from the first operator pass, sum_i sq_i^2 [u~ (E + X(z)) u~ + u~ h u~ - (h u~) o (h u~)]_i,
which equals sum(u w) on the symmetric subspace; the coefficients agree with the oracle to the parity bar (1e-9) and with
the separate pass to rounding, with and without the projection, with members that leave at their own depth, across the
rescaling of the lagged normalisation (every 16 levels), with o = 0 and with e != 0, at member tiles 8 and 16.

The chain blocks take the same fused level: route ed3 (chain, src/asr_ed3.f90:169-184; a_n from the first pass as
sum_i sq_i^2 [u~ (z - et) u~ + u~ h1 - h1 (u u~) - h1 t]_i with h1 = h' u~, t = u u~ - Qt h1) and the EPI pair blocks
(the edc operator), both with the projections c_n = <t, u_n> of the second vector from the update inside the second pass:
a_n, b_n, c_n and the blocks against the separate pass to rounding and against the oracle to the parity bar.
"""
import numpy as np
import pytest

from asr_amd import workloads as W

pytestmark = pytest.mark.gpu

KIND_UPDATE, KIND_SPMM_UPD = 3, 7


def _engine():
    from asr_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    return engine


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300))


@pytest.fixture(scope="module")
def wl():
    w = W.ternary_edc(L=4, levels=2, depth=20, seed=12)
    rng = np.random.default_rng(8)
    members = list(w.members)
    # members with e != 0, and members with o = 0 (then the level has no o pass)
    for k in (3, 20, 41):
        m = W.edc_member(w.info["params"][k // 27], (k // 9) % 3, k % 9, w.members[k].z)
        m.e = rng.uniform(-0.1, 0.1, m.e.shape)
        members[k] = m
    return w, members


def _oracle(op):
    from oracle.oracle import OracleOp
    return OracleOp(op)


@pytest.mark.parametrize("tile", [8, 16])
@pytest.mark.parametrize("project", [True, False])
def test_fused_level_against_oracle_and_separate_pass(wl, tile, project):
    eng = _engine()
    w, members = wl
    nlev = np.array([20 - (k % 7) * 2 for k in range(len(members))], np.int32)  # 8..20 levels, past the rescale at 16
    G = eng.GpuOperator(w.op)
    G.set_tile(tile)
    out = {}
    for form in (0, 1):
        G.set_level_form(form)
        out[form] = G.chains_edc(members, nlev, project=project)
    a0, b0, mu0, _ = out[0]
    a1, b1, mu1, _ = out[1]
    assert np.array_equal(mu0, mu1)
    assert rel(a0, a1) < 1e-12 and rel(b0, b1) < 1e-12
    O = _oracle(w.op)
    for j in (0, 3, 20, 26, 41, 53):
        m = members[j]
        ar, br, n = O.edc_chain(m, m.z * m.pa + m.pb, m.l0, int(nlev[j]), project=project)
        assert mu0[j] == n
        assert rel(a0[j, :n], ar[:n]) < 1e-9, f"chain {j}"
        assert rel(b0[j, :n - 1], br[:n - 1]) < 1e-9, f"chain {j}"


def test_fused_level_zero_o(wl):
    eng = _engine()
    w, members = wl
    sel = [members[k] for k in (0, 9, 30, 44)]
    for m in sel:
        m.o = np.zeros_like(m.o)
    G = eng.GpuOperator(w.op)
    G.set_tile(8)
    a, b, mu, _ = G.chains_edc(sel, 18)
    O = _oracle(w.op)
    for j, m in enumerate(sel):
        ar, br, n = O.edc_chain(m, m.z * m.pa + m.pb, m.l0, 18)
        assert mu[j] == n and rel(a[j, :n], ar[:n]) < 1e-9 and rel(b[j, :n - 1], br[:n - 1]) < 1e-9


def test_fused_level_is_the_path_taken(wl):
    """The level's launches: form 0 has the update inside the second operator pass (kind 7) and no update pass (kind 3);
    form 1 the reverse."""
    eng = _engine()
    w, members = wl
    G = eng.GpuOperator(w.op)
    G.set_tile(8)
    G.set_level_form(0)
    kinds0 = list(G.profile(members, 6)[1])
    G.set_level_form(1)
    kinds1 = list(G.profile(members, 6)[1])
    assert KIND_SPMM_UPD in kinds0 and KIND_UPDATE not in kinds0, kinds0
    assert KIND_UPDATE in kinds1 and KIND_SPMM_UPD not in kinds1, kinds1


def test_fused_level_graph_and_reproducible(wl):
    """One captured HIP graph (chains_multi mode 2) equals the host-issued levels bit for bit, run after run."""
    eng = _engine()
    w, members = wl
    G = eng.GpuOperator(w.op)
    G.set_tile(8)
    a, b, mu, _ = G.chains_edc(members, 20)
    a2, b2, mu2, _ = G.chains_edc(members, 20)
    assert np.array_equal(a, a2) and np.array_equal(b, b2)
    res, _ = eng.chains_multi([(G, members, np.full(len(members), 20, np.int32))], mode=2)
    assert np.array_equal(res[0][0], a) and np.array_equal(res[0][1], b)


def _asymmetric_operator():
    """Bond blocks chosen by the type of the node's own site only (type 1: x 1.1): S is not symmetric."""
    from asr_amd import augcls as A
    from asr_amd import structure as S
    g, block, block0 = W.graph_and_blocks("fcc", 4)
    cls = np.zeros(g.n, np.int32)
    perm, mats = S.class_ops(g.coord, 8, cls, "d2h")
    fl = np.ones(g.n, np.uint8)
    fl[0] = 0
    stype = (np.arange(g.n) % 2).astype(np.int32)
    return A.reorder_nodes(A.build_operator(g, cls, [3], [[0, 1, 2]], [[0.66, 0.14, 0.2]], perm, S.orbital_rep(mats), 0, 2,
                                            maxrank=1, flippable=fl, stype=stype, block=np.stack([block, 1.1 * block]),
                                            block0=np.stack([block0, block0]), nsp=3))


def test_fused_level_falls_back_on_an_asymmetric_operator():
    """Bond blocks chosen by the type of the node's own site only (type 1: x 1.1) make S non-symmetric; the fused level's
    a_n then differs from the direct <u, H u>, the run is repeated with the separate update pass (the oracle's chains),
    and the operator keeps that form."""
    eng = _engine()
    op = _asymmetric_operator()
    spins = W.edc_params(np.random.default_rng(5), 3, 1)
    members = [W.edc_member(spins[0], cen, l0, complex(-0.2, 0.15)) for cen in range(3) for l0 in (0, 4, 8)]
    G = eng.GpuOperator(op)
    G.set_tile(8)
    a, b, mu, _ = G.chains_edc(members, 10)
    O = _oracle(op)
    for j in (0, 4, 8):
        m = members[j]
        ar, br, n = O.edc_chain(m, m.z * m.pa + m.pb, m.l0, 10)
        assert mu[j] == n and rel(a[j, :n], ar[:n]) < 1e-9 and rel(b[j, :n - 1], br[:n - 1]) < 1e-9
    kinds = list(G.profile(members, 4)[1])
    assert KIND_UPDATE in kinds and KIND_SPMM_UPD not in kinds, kinds


def _block_forms(G, run):
    out = {}
    for form in (0, 1):
        G.set_level_form(form)
        out[form] = run()
    G.set_level_form(0)
    return out


@pytest.mark.parametrize("tile", [8, 16])
def test_fused_level_ed3_blocks(tile):
    eng = _engine()
    from asr_amd import ed3 as E3
    from oracle.oracle import OracleOp
    wl = W.ternary_ed3(L=4, levels=2, depth=20, seed=21)
    sel = list(range(0, len(wl.members), 2))
    members = [wl.members[j] for j in sel]
    nlev = np.array([20 - (k % 5) * 3 for k in range(len(sel))], np.int32)  # 8..20 levels, past the rescale at 16
    G = eng.GpuOperator(wl.op)
    G.set_tile(tile)
    G.set_depth_tol(0.0, "hull")
    out = _block_forms(G, lambda: G.ed3_blocks(members, wl.info["cf"][sel], wl.info["oc"][sel], nlev, mode=0))
    (bk0, a0, b0, c0, mu0, _), (bk1, a1, b1, c1, mu1, _) = out[0], out[1]
    assert np.array_equal(mu0, mu1)
    # the fused a_n is a different sum: equal to rounding, not bit for bit (so the fused level did run)
    assert not np.array_equal(a0, a1)
    # c_n falls to 1e-5 of c_0 over the chain: its rounding, relative to the largest, reaches a few 1e-12; the blocks take
    # the hull tail's eigenvalues of the chain matrix, which amplify the coefficients' rounding (the oracle bar is 1e-8)
    assert rel(a0, a1) < 1e-12 and rel(b0, b1) < 1e-12 and rel(c0, c1) < 1e-10 and rel(bk0, bk1) < 1e-9
    O = OracleOp(wl.op)
    for q in (0, 3, len(sel) - 1):
        j = sel[q]
        m = wl.members[j]
        et, uu, qt = E3.ed3_coefs(wl.info["ed"][j], m.z)
        bko, chains, _ = O.ed3_blocks(m, et, uu, qt, wl.info["oc"][j], m.z, m.l0, int(nlev[q]), project=True, tail="hull")
        for sd in range(2):
            ao, bo, co, mo, closed = chains[sd]
            assert mu0[q, sd] == mo
            assert rel(a0[q, sd, :mo], ao[:mo]) < 1e-9 and rel(b0[q, sd, :mo - 1], bo[:mo - 1]) < 1e-9, (j, sd)
            assert rel(c0[q, sd, :mo], co[:mo]) < 1e-9, (j, sd)
        assert rel(bk0[q], bko) < 1e-8, j


@pytest.mark.parametrize("tile", [8, 16])
def test_fused_level_epi_blocks(tile):
    eng = _engine()
    from oracle.oracle import OracleOp
    wl = W.fcc_epi_pair(L=4, levels=2, seed=31)
    nlev = np.array([20 - (k % 5) * 3 for k in range(len(wl.members))], np.int32)
    G = eng.GpuOperator(wl.op)
    G.set_tile(tile)
    G.set_depth_tol(0.0, "hull")
    out = _block_forms(G, lambda: G.epi_blocks(wl.members, wl.info["v9"], nlev, mode=0))
    (x0, a0, b0, c0, mu0, _), (x1, a1, b1, c1, mu1, _) = out[0], out[1]
    assert np.array_equal(mu0, mu1)
    assert not np.array_equal(a0, a1)
    assert rel(a0, a1) < 1e-12 and rel(b0, b1) < 1e-12 and rel(c0, c1) < 1e-10 and rel(x0, x1) < 1e-9
    O = OracleOp(wl.op)
    for j in (0, 7, len(wl.members) - 1):
        m = wl.members[j]
        xo, chains = O.epi_blocks(m, m.z * m.pa + m.pb, wl.info["v9"][j], m.z, int(nlev[j]), project=True, tail="hull")
        for sd in range(2):
            ao, bo, co, mo, closed = chains[sd]
            assert mu0[j, sd] == mo
            assert rel(a0[j, sd, :mo], ao[:mo]) < 1e-9 and rel(b0[j, sd, :mo - 1], bo[:mo - 1]) < 1e-9, (j, sd)
            assert rel(c0[j, sd, :mo], co[:mo]) < 1e-9, (j, sd)
        assert rel(x0[j], xo) < 1e-8, j


def test_fused_block_level_falls_back_on_an_asymmetric_operator():
    """Route ed3's blocks on the same non-symmetric operator: the fused level's a_n differs from the direct one, the run
    is repeated with the separate update pass, and the chains equal the oracle's."""
    eng = _engine()
    from asr_amd import ed3 as E3
    from oracle.oracle import OracleOp
    wl = W.ternary_ed3(L=4, levels=2, depth=10, seed=21, op=_asymmetric_operator())
    sel = [0, 4, 8, 31]
    members = [wl.members[j] for j in sel]
    G = eng.GpuOperator(wl.op)
    G.set_tile(8)
    G.set_depth_tol(0.0, "hull")
    bk, a, b, c, mu, _ = G.ed3_blocks(members, wl.info["cf"][sel], wl.info["oc"][sel], 10, mode=0)
    O = OracleOp(wl.op)
    for q, j in enumerate(sel):
        m = wl.members[j]
        et, uu, qt = E3.ed3_coefs(wl.info["ed"][j], m.z)
        bko, chains, _ = O.ed3_blocks(m, et, uu, qt, wl.info["oc"][j], m.z, m.l0, 10, project=True, tail="hull")
        for sd in range(2):
            ao, bo, co, mo, closed = chains[sd]
            assert mu[q, sd] == mo
            assert rel(a[q, sd, :mo], ao[:mo]) < 1e-9 and rel(b[q, sd, :mo - 1], bo[:mo - 1]) < 1e-9, (j, sd)
            assert rel(c[q, sd, :mo], co[:mo]) < 1e-9, (j, sd)
