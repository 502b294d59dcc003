"""GPU parity on the reference's own point group and the reduced-against-unreduced identity on the device.

* D4h (class_ops default, src/asr_symmetry.f90:67-92) permutes p_x / p_y and d_xz / d_yz, so pi is not the identity:
  the operator passes take their orbital-permuting forms (k_tile<.., PI = false> at member tiles 8 and 16, with the fused
  level's projection over the stabiliser in registers; k_fused at 32, k_node<T, false> at 64, k_pair<false> at 65, with
  the update's general projector).  Chains and applications against the oracle at tiles 8, 16, 32, 64 and 65; the fused
  level against the separate update pass.
* On the GPU alone: the D4h- and D2h-reduced chains equal the chains of the unreduced operator (identity group), and the
  144-step chain of the reduced ordered B2 cell equals the dense solve on the unreduced one, 1e-11
  (test/test_augcls_project.f90:100-150).
* Vectors of 4 GiB and more take the pointer forms of the operator pass (the buffer-load forms use 32-bit offsets):
  512 real chains at tile 64 and 270 complex chains at tile 65 on the 127,517-node operator, against the oracle.
"""
import numpy as np
import pytest

from asr_amd import workloads as W

import reduced as R

pytestmark = pytest.mark.gpu


def _engine():
    from asr_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    return engine


def _oracle(op):
    from oracle.oracle import OracleOp
    return OracleOp(op)


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300))


@pytest.fixture(scope="module")
def d4h_ternary():
    wl = W.ternary_edc(L=4, levels=2, depth=14, seed=12, group="d4h")
    assert not np.all(wl.op.pi == np.arange(9)[None, :]), "D4h must permute orbitals"
    members = [m for m in wl.members if m.l0 in R.L0S_1D]  # 30 chains: 2 spins x 3 species x 5 orbitals
    return wl, members


@pytest.mark.parametrize("tile", [8, 16, 32, 64, 65])
def test_d4h_edc_chains_every_tile(d4h_ternary, tile):
    eng = _engine()
    wl, members = d4h_ternary
    print(f"d4h edc: {wl.op.nn} nodes, tile {tile}, {len(members)} chains", flush=True)
    G = eng.GpuOperator(wl.op)
    G.set_tile(tile)
    nlev = np.array([8 + (k % 5) for k in range(len(members))], np.int32)
    for project in (True, False):
        a, b, mu, _ = G.chains_edc(members, nlev, project=project)
        O = _oracle(wl.op)
        for j in (0, 7, 16, len(members) - 1):
            m = members[j]
            ar, br, n = O.edc_chain(m, m.z * m.pa + m.pb, m.l0, int(nlev[j]), project=project)
            assert mu[j] == n
            assert rel(a[j, :n], ar[:n]) < 1e-9, f"chain {j} project {project}"
            assert rel(b[j, :n - 1], br[:n - 1]) < 1e-9, f"chain {j} project {project}"


@pytest.mark.parametrize("tile", [8, 16])
def test_d4h_fused_level_against_separate_pass(d4h_ternary, tile):
    """The fused level on the orbital-permuting group (the stabiliser projection sum_g chi sg w_pi(g, l) / |S| on the
    registers of the second operator pass) equals the separate update pass (k_update's general projector) to rounding,
    and is the path taken."""
    eng = _engine()
    wl, members = d4h_ternary
    G = eng.GpuOperator(wl.op)
    G.set_tile(tile)
    out = {}
    for form in (0, 1):
        G.set_level_form(form)
        out[form] = G.chains_edc(members, 14)
    assert np.array_equal(out[0][2], out[1][2])
    assert rel(out[0][0], out[1][0]) < 1e-12 and rel(out[0][1], out[1][1]) < 1e-12
    G.set_level_form(0)
    kinds = list(G.profile(members, 4)[1])
    assert 7 in kinds and 3 not in kinds, kinds


@pytest.mark.parametrize("tile", [8, 16, 64])
def test_d4h_real_chains_every_tile(tile):
    eng = _engine()
    wl = W.plain_binary(L=4, levels=2, depth=12, seed=11, group="d4h")
    members = [m for m in wl.members if m.l0 in R.L0S_1D] * 3
    print(f"d4h lanczos: {wl.op.nn} nodes, tile {tile}, {len(members)} chains", flush=True)
    G = eng.GpuOperator(wl.op)
    G.set_tile(tile)
    a, b, mu, _ = G.lanczos(members, 12)
    O = _oracle(wl.op)
    for j in (0, 3, 9, len(members) - 1):
        ar, br, n = O.lanczos(members[j], members[j].l0, 12)
        assert mu[j] == n
        assert rel(a[j, :n], ar[:n]) < 1e-9 and rel(b[j, :n], br[:n]) < 1e-9, f"chain {j}"


def test_d4h_apply_real_and_complex(d4h_ternary):
    eng = _engine()
    wl, members = d4h_ternary
    op = wl.op
    rng = np.random.default_rng(2)
    sel = [members[k] for k in (0, 6, 13, 29)]
    G = eng.GpuOperator(op)
    O = _oracle(op)
    x = rng.standard_normal((4, op.dim))
    y = G.apply(sel, x)
    for j, m in enumerate(sel):
        assert rel(y[j], O.apply(m, x[j])) < 1e-12, f"member {j}"
    xc = x + 1j * rng.standard_normal((4, op.dim))
    yc = G.apply_complex(sel, xc)
    for j, m in enumerate(sel):
        ep = m.z * m.pa + m.pb
        ref = O.apply(m, xc[j].real) + 1j * O.apply(m, xc[j].imag) + O.rand_c(m, ep, xc[j])
        assert rel(yc[j], ref) < 1e-12, f"member {j}"


@pytest.mark.parametrize("group", ["d2h", "d4h"])
def test_gpu_reduced_equals_unreduced(group):
    """The same chains on the reduced and on the unreduced node set (20,455 nodes), both on the GPU: 1e-9."""
    eng = _engine()
    wr = W.ternary_edc(L=4, levels=2, depth=12, seed=12, group=group)
    wu = W.ternary_edc(L=4, levels=2, depth=12, seed=12, group="identity")
    l0s = list(range(9)) if group == "d2h" else R.L0S_1D
    members = [m for m in wr.members if m.l0 in l0s]
    print(f"{group}: {wr.op.nn} reduced nodes, {wu.op.nn} unreduced, {len(members)} chains", flush=True)
    Gr, Gu = eng.GpuOperator(wr.op), eng.GpuOperator(wu.op)
    ar, br, mr, _ = Gr.chains_edc(members, 12)
    au, bu, mu, _ = Gu.chains_edc(members, 12)
    assert np.array_equal(mr, mu)
    assert rel(ar, au) < 1e-9 and rel(br[:, :11], bu[:, :11]) < 1e-9
    ar, br, mr, _ = Gr.lanczos(members, 12)
    au, bu, mu, _ = Gu.lanczos(members, 12)
    assert rel(ar, au) < 1e-9 and rel(br, bu) < 1e-9


def test_gpu_b2_reduced_chain_equals_dense_unreduced_solve():
    from test_reduction import dense_parts, g_none
    eng = _engine()
    opr, opu = R.b2_operator("d4h"), R.b2_operator("identity")
    pp = R.pot_params(2, 20261110)
    z = complex(0.15, 0.03)
    members = [R.member(pp, 0, l0, z) for l0 in R.L0S_1D]
    G = eng.GpuOperator(opr)
    for tile in (8, 64, 65):
        G.set_tile(tile)
        a, b, mu, _ = G.chains_edc(members, 144)
        Ou = _oracle(opu)
        for j, m in enumerate(members):
            h, xa, xb = dense_parts(Ou, m)
            k = z * np.eye(opu.dim) - h - (z * xa + xb)
            rhs = np.zeros(opu.dim, complex)
            rhs[m.l0] = 1.0
            gex = np.linalg.solve(k, rhs)[m.l0]
            g = g_none(a[j], b[j], int(mu[j]), z)
            assert abs(g - gex) / abs(gex) < 1e-11, f"tile {tile} orbital {m.l0}"


# ------------------------------------------------------------------------------------------------ vectors of 4 GiB and more
@pytest.fixture(scope="module")
def headline_adjoint():
    return W.ternary_adjoint(L=8, levels=4, nchains=512, depth=8, seed=5)


def test_real_tile64_beyond_4gib(headline_adjoint):
    """512 real members at tile 64 on the 127,517-node operator: 4.7 GB per vector, beyond the 32-bit buffer offsets of
    the buffer-load S pass; the pass takes its pointer form and gives the oracle's chains."""
    eng = _engine()
    wl = headline_adjoint
    nv = (wl.op.nn + 1) * 9 * 512 * 8
    assert nv >= 2 ** 32
    print(f"real tile 64: {wl.op.nn} nodes, 512 chains, {nv / 1e9:.2f} GB per vector", flush=True)
    G = eng.GpuOperator(wl.op)
    G.set_tile(64)
    a, b, mu, _ = G.lanczos(wl.members, 8)
    G.close()
    O = _oracle(wl.op)
    for j in (0, 511):
        m = wl.members[j]
        ar, br, n = O.lanczos(m, m.l0, 8)
        assert mu[j] == n == 8
        assert rel(a[j], ar) < 1e-9 and rel(b[j], br) < 1e-9, f"chain {j}"


def test_complex_tile65_beyond_4gib():
    """270 complex members at tile 65 (lane pairs, 320 padded) on the 127,517-node operator: 5.9 GB per vector."""
    eng = _engine()
    wl = W.ternary_edc(L=8, levels=4, depth=6, seed=2)
    members = wl.members * 5
    nv = (wl.op.nn + 1) * 9 * 320 * 16
    assert nv >= 2 ** 32
    print(f"complex tile 65: {wl.op.nn} nodes, {len(members)} chains, {nv / 1e9:.2f} GB per vector", flush=True)
    G = eng.GpuOperator(wl.op)
    G.set_tile(65)
    a, b, mu, _ = G.chains_edc(members, 6)
    G.close()
    O = _oracle(wl.op)
    for j in (0, 269):
        m = members[j]
        ar, br, n = O.edc_chain(m, m.z * m.pa + m.pb, m.l0, 6)
        assert mu[j] == n
        assert rel(a[j, :n], ar[:n]) < 1e-9 and rel(b[j, :n - 1], br[:n - 1]) < 1e-9, f"chain {j}"
