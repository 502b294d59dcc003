"""Pins the oracle (and the node-set builder that feeds it) to independent truths, on the CPU.

* Ring brute force (test_slow/test_augcls.f90:82-148, test/test_augn.f90:41-100): recursion with an exhausted Krylov space
  equals the exact configuration average of the resolvent, for three classes (ordered / binary / ternary), the start site
  frozen to each allowed species, real (energy-independent) and complex-symmetric (energy-dependent) chains.
* Symmetry of the matrix-free operator (test/test_augn.f90:46-49).
* Second moment a_1^2 + b_1^2 = <u|H^2|u> (test/test_recursion_dense.f90:101-105).
"""
import numpy as np
import pytest

from asr_amd import augcls as A
from asr_amd import host
from oracle.oracle import OracleOp

from exact import build_ring_operator, dense_operator, exact_average_asa, exact_average_g, exact_average_pair_blocks, ring_model


def rotated(model, ic):
    ib = int(np.nonzero(model["cls"] == ic)[0][0])
    m = dict(model)
    m["cls"] = np.roll(model["cls"], -ib)
    return m


CASES = [(ic, cen) for ic, spc in enumerate([[0], [1, 2], [0, 2, 3]]) for cen in spc]


@pytest.mark.parametrize("ic,cen", CASES)
def test_ring_real_chain_equals_exact_average(ic, cen):
    model, par = ring_model()
    model = rotated(model, ic)
    op = build_ring_operator(model, frozen_start=True)
    n_expect = 4 * int(np.prod([model["nsc"][k] for k in model["cls"][1:]]))
    assert op.nn == n_expect
    O = OracleOp(op)
    l0 = 4
    m = A.Member(l0=l0, c=par["c"], d=par["d"], o=par["o"], e=par["e"], frozen=cen)
    a, b, n = O.lanczos(m, l0, op.dim, tiny=1.0e-9)
    err = 0.0
    for e in (-0.6, 0.0, 0.6):
        z = complex(e, 0.3)
        g = host.terminated_green(a, b, n, 0.0, 0.0, z)[0]
        gx = exact_average_g(model, par, z, l0, cen)
        err = max(err, abs(g - gx))
    assert err < 1.0e-7


@pytest.mark.parametrize("ic,cen", CASES[:4])
def test_ring_edc_operator_and_chain(ic, cen):
    """The energy-dependent operator K = H + X(z) (augcls_apply on Re/Im + augcls_rand_c) has the exact configuration
    average as its resolvent (dense solve, 1e-10), and the complex-symmetric chain reproduces that solve (1e-6)."""
    model, par = ring_model()
    model = rotated(model, ic)
    op = build_ring_operator(model, frozen_start=True)
    O = OracleOp(op)
    mem = A.Member(l0=0, c=par["c"], d=par["d"], o=par["o"], e=np.zeros((4, 9)), frozen=cen)
    dim = op.dim
    err_op = err_ch = 0.0
    for e in (-0.3, 0.3):
        z = complex(e, 0.5)
        ep = z * par["pa"] + par["pb"]
        k = np.zeros((dim, dim), np.complex128)
        for i in range(dim):
            u = np.zeros(dim)
            u[i] = 1.0
            k[:, i] = O.apply(mem, u) + O.rand_c(mem, ep, u.astype(np.complex128))
        rhs = np.zeros(dim, np.complex128)
        rhs[0] = 1.0
        gaug = np.linalg.solve(z * np.eye(dim) - k, rhs)[0]
        gx = exact_average_g(model, par, z, 0, cen, ep=True)
        err_op = max(err_op, abs(gaug - gx))
        a, b, mm = O.edc_chain(mem, ep, 0, 1500, project=True)  # no reorthogonalisation: about 3x dim steps here
        g, _ = host.cf_species(a, b, np.zeros(mm), mm, z, tail="none")
        err_ch = max(err_ch, abs(g - gaug))
    assert err_op < 1.0e-10
    assert err_ch < 1.0e-6


def test_operator_symmetric_and_second_moment():
    model, par = ring_model()
    op = build_ring_operator(rotated(model, 2), frozen_start=False, maxlevel=80)
    O = OracleOp(op)
    m = A.Member(l0=0, c=par["c"], d=par["d"], o=par["o"], e=par["e"])
    h = dense_operator(lambda x: O.apply(m, x), op.dim)
    assert np.max(np.abs(h - h.T)) < 1.0e-12 * np.max(np.abs(h))
    a, b, n = O.lanczos(m, 0, 6)
    u = np.zeros(op.dim)
    u[0] = 1.0
    hu = h @ u
    assert abs(a[0] ** 2 + b[0] ** 2 - hu @ hu) < 1.0e-12 * (hu @ hu)


def test_oracle_mutation_is_detected():
    """A planted mistake (one sign flipped in one bond block) must break the exact-average agreement."""
    model, par = ring_model()
    model = rotated(model, 1)
    op = build_ring_operator(model, frozen_start=True)
    op.block = op.block.copy()
    op.block[0, 0, 2, 5] = -op.block[0, 0, 2, 5]
    O = OracleOp(op)
    m = A.Member(l0=4, c=par["c"], d=par["d"], o=par["o"], e=par["e"], frozen=1)
    a, b, n = O.lanczos(m, 4, op.dim, tiny=1.0e-9)
    z = complex(0.0, 0.3)
    g = host.terminated_green(a, b, n, 0.0, 0.0, z)[0]
    assert abs(g - exact_average_g(model, par, z, 4, 1)) > 1.0e-6


def test_node_reordering_keeps_the_chains():
    """The locality relabelling of the node set changes no chain coefficient beyond rounding (oracle on both orders)."""
    from asr_amd import workloads as W
    w0 = W.ternary_edc(L=4, levels=2, depth=10, reorder=False)
    w1 = W.ternary_edc(L=4, levels=2, depth=10, reorder=True)
    assert w1.op.meta.get("node_order") == "rcm of the hop links" and not np.array_equal(w0.op.hopr, w1.op.hopr)
    O0, O1 = OracleOp(w0.op), OracleOp(w1.op)
    for k in (0, 17, 40):
        m = w0.members[k]
        ep = m.z * m.pa + m.pb
        a0, b0, n0 = O0.edc_chain(m, ep, m.l0, 10)
        a1, b1, n1 = O1.edc_chain(m, ep, m.l0, 10)
        assert n0 == n1
        assert np.max(np.abs(a1 - a0)) < 1e-12 * np.max(np.abs(a0))
        assert np.max(np.abs(b1 - b0)) < 1e-12 * np.max(np.abs(b0))


def ring_ed3_params(seed: int = 77):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.1, 0.1, (4, 3))
    return dict(c=rng.uniform(-0.4, 0.4, (4, 3)), sqrdel=rng.uniform(0.5, 0.8, (4, 3)), enu=rng.uniform(-0.1, 0.1, (4, 3)),
                o=o, p=o ** 2 + rng.uniform(0.0, 0.05, (4, 3)))


@pytest.mark.parametrize("ic,cen", CASES[:4])
def test_ring_ed3_blocks_equal_exact_average(ic, cen):
    """Route ed3 on the ring (test/test_asr_ed3.f90): the augmented K'(z) is complex symmetric; its dense inverse gives the
    configuration averages of G_11, (hG)_11, (hGh)_11 of K = E O - H (1e-10); the two projected chains of ed3_blocks
    (exhausted Krylov space, no tail) give them too (1e-6), and the two cross terms agree."""
    from asr_amd import ed3 as E3
    model, _ = ring_model()
    model = rotated(model, ic)
    op = build_ring_operator(model, frozen_start=True)
    O = OracleOp(op)
    pp = ring_ed3_params()
    ed = E3.ed3_setup(pp, cen)
    orb = {k: E3.per_orbital(pp[k]) for k in ("c", "enu", "o", "p")}
    orb["dl"] = E3.per_orbital(pp["sqrdel"]) ** 2
    dim = op.dim
    err_sym = err_dense = err_chain = err_cross = 0.0
    for l0 in (0, 2, 6):
        for e in (-0.3, 0.25):
            z = complex(e, 0.3)
            m = E3.ed3_member(ed, l0, z)
            et, uu, qt = E3.ed3_coefs(ed, z)
            k = np.zeros((dim, dim), np.complex128)
            for j in range(dim):
                u = np.zeros(dim, np.complex128)
                u[j] = 1.0
                k[:, j] = O.ed3_apply(m, et, uu, qt, u)
            err_sym = max(err_sym, np.max(np.abs(k - k.T)) / np.max(np.abs(k)))
            vv, av, nv, na = O.ed3_start(m, l0, E3.ed3_oc(ed, l0), project=True)
            rhs = np.zeros((dim, 2), np.complex128)
            rhs[l0, 0] = 1.0
            rhs[:, 1] = vv
            sol = np.linalg.solve(k, rhs)
            dense = np.array([sol[l0, 0], vv @ sol[:, 0], vv @ sol[:, 1]])
            exact = np.array(exact_average_asa(model, orb, z, l0, cen))
            err_dense = max(err_dense, np.max(np.abs(dense - exact)))
            bk, chains, _ = O.ed3_blocks(m, et, uu, qt, E3.ed3_oc(ed, l0), z, l0, 1500, project=True, tail="none")
            err_chain = max(err_chain, np.max(np.abs(bk[:3] - exact)))
            err_cross = max(err_cross, abs(bk[3]))
    assert err_sym < 1.0e-13
    assert err_dense < 1.0e-10
    assert err_chain < 1.0e-6
    assert err_cross < 1.0e-6


def test_ed3_mutation_is_detected():
    """A planted mistake in the oracle's K' (the sign of the Qt term's coefficient) breaks the exact-average agreement."""
    from asr_amd import ed3 as E3
    model, _ = ring_model()
    model = rotated(model, 1)
    op = build_ring_operator(model, frozen_start=True)
    O = OracleOp(op)
    pp = ring_ed3_params()
    ed = E3.ed3_setup(pp, 1)
    orb = {k: E3.per_orbital(pp[k]) for k in ("c", "enu", "o", "p")}
    orb["dl"] = E3.per_orbital(pp["sqrdel"]) ** 2
    z = complex(0.1, 0.3)
    m = E3.ed3_member(ed, 2, z)
    et, uu, qt = E3.ed3_coefs(ed, z)
    bk, _, _ = O.ed3_blocks(m, et, uu, -qt, E3.ed3_oc(ed, 2), z, 2, 1500, project=True, tail="none")
    exact = np.array(exact_average_asa(model, orb, z, 2, 1))
    assert np.max(np.abs(bk[:3] - exact)) > 1.0e-4


def test_lanczos_full_reorthogonalisation():
    """lanczos with reorth = .true. (src/reference/asr_recursion.f90:17-58): with every stored vector projected out twice
    the chain spans the whole space here (rounding splits the degenerate pairs of the ring's reflection) and closes
    (b_n < tiny) after dim steps; the chain matrix then has exactly the spectrum of the dense operator, multiplicities
    included, which a chain without reorthogonalisation (ghost copies) does not; the Green function is the exact
    configuration average to 1e-10."""
    model, par = ring_model()
    model = rotated(model, 2)
    op = build_ring_operator(model, frozen_start=True)
    O = OracleOp(op)
    m = A.Member(l0=4, c=par["c"], d=par["d"], o=par["o"], e=par["e"], frozen=0)
    h = dense_operator(lambda x: O.apply(m, x), op.dim)
    ev = np.linalg.eigvalsh(0.5 * (h + h.T))
    a, b, n = O.lanczos(m, 4, op.dim + 8, tiny=1.0e-9, reorth=True)
    assert n == op.dim and b[n - 1] < 1.0e-9
    ritz = host.tridiag_eigvals(a[:n], b[:n - 1])
    assert np.max(np.abs(np.sort(ritz) - ev)) < 1.0e-9
    a0, b0, n0 = O.lanczos(m, 4, op.dim, tiny=1.0e-9)
    ritz0 = host.tridiag_eigvals(a0[:n0], b0[:n0 - 1])
    assert np.max(np.abs(np.sort(ritz0) - ev)) > 1.0e-3
    for e in (-0.5, 0.2):
        z = complex(e, 0.3)
        g = host.terminated_green(a, b, n, 0.0, 0.0, z)[0]
        assert abs(g - exact_average_g(model, par, z, 4, 0)) < 1.0e-10


def _ring_pair_setup(seed: int = 91):
    """Ring with site 0 (class {1, 2}) on a path between species 1 and 2 and the pair partner site 2 (class {1, 2}) frozen
    to species 2; site 1 (ternary class) random; the path point is species 4."""
    from asr_amd import epi as EP
    from exact import ring_graph
    model, _ = ring_model()
    model = rotated(model, 1)
    partner = 2
    g = ring_graph(4)
    fl = np.ones(4, np.uint8)
    fl[0] = 0
    fl[partner] = 0
    op = A.build_operator(g, model["cls"], model["nsc"], model["spc"], model["conc"], np.arange(4, dtype=np.int32)[None, :],
                          np.eye(9)[None], 0, 80, flippable=fl, block=model["block"], block0=model["block0"], nsp=5,
                          partner=partner)
    rng = np.random.default_rng(seed)
    par = [EP.epi_species(dict(enu=rng.uniform(-0.1, 0.1, 3), c=rng.uniform(-0.4, 0.4, 3), sqrdel=rng.uniform(0.5, 0.8, 3),
                               o=rng.uniform(-0.1, 0.1, 3))) for _ in range(4)]
    pt = EP.epi_path(par[1], par[2], 0.37)
    sites = {k: np.stack([getattr(p, k) for p in par] + [getattr(pt, k)]) for k in ("c", "d", "o", "e")}
    return model, op, par, pt, sites, partner


def test_ring_epi_blocks_equal_exact_average():
    """The EPI blocks (epi_blocks, src/asr_epi.f90:628-647) on the ring with two frozen sites (site 0 on the path point,
    the partner frozen to a species: augcls_freeze_any): the dense resolvent of the scaled augmented operator gives the
    conditional average of X0 = v9^T G_00 v9 of the unscaled H = E_nu + h - h o h (1e-10), and the two projected chains
    give X0, X1 = (hG)_00, X2 = (hGh)_00 (1e-6, exhausted Krylov space) with a vanishing cross term, for a random unit
    start combination v9 (the identity group: chi = 1)."""
    from asr_amd import epi as EP
    model, op, par, pt, sites, partner = _ring_pair_setup()
    O = OracleOp(op)
    rng = np.random.default_rng(5)
    assert set(np.unique(op.fz)) == {0, 1, 2}
    err_dense = err_chain = err_cross = 0.0
    for v9 in (np.eye(9)[4], rng.standard_normal(9)):
        v9 = v9 / np.linalg.norm(v9)
        for e in (-0.25, 0.2):
            z = complex(e, 0.3)
            m = EP.epi_member(par, pt, 4, np.ones(1), 2, z)
            ep = z * m.pa + m.pb
            dim = op.dim
            k = np.zeros((dim, dim), np.complex128)
            for j in range(dim):
                u = np.zeros(dim)
                u[j] = 1.0
                k[:, j] = O.apply(m, u) + O.rand_c(m, ep, u.astype(np.complex128))
            s0 = np.zeros(dim)
            s0[:9] = v9
            x0 = s0 @ np.linalg.solve(z * np.eye(dim) - k, s0)
            exact = exact_average_pair_blocks(model, sites, z, v9, partner, 2, 4)
            err_dense = max(err_dense, abs(x0 - exact[0]))
            x, chains = O.epi_blocks(m, ep, v9, z, 1500, tail="none")
            err_chain = max(err_chain, np.max(np.abs(x[:3] - exact)))
            err_cross = max(err_cross, abs(x[3]))
    assert err_dense < 1.0e-10
    assert err_chain < 1.0e-6
    assert err_cross < 1.0e-6
    # a planted mistake (the partner frozen to the other species of its class) is detected
    m = EP.epi_member(par, pt, 4, np.ones(1), 1, z)
    x, _ = O.epi_blocks(m, z * m.pa + m.pb, v9, z, 1500, tail="none")
    assert np.max(np.abs(x[:3] - exact)) > 1.0e-4


def test_epi_adapted_vectors():
    """epi_adapted (src/asr_epi.f90:290-328): for the abelian D2h the nine vectors are orthonormal, one l shell each, and
    common eigenvectors of every operation with characters +-1; with the identity alone they are the harmonics."""
    from asr_amd import epi as EP
    from asr_amd import structure as S
    g = S.periodic_graph("fcc", 2)
    perm, mats = S.class_ops(g.coord, 4, np.zeros(g.n, np.int32), "d2h")
    d = S.orbital_rep(mats)
    vec, chi, lo = EP.epi_adapted(d)
    assert np.allclose(vec.T @ vec, np.eye(9), atol=1e-12)
    for a in range(9):
        for gq in range(d.shape[0]):
            assert np.allclose(d[gq] @ vec[:, a], chi[gq, a] * vec[:, a], atol=1e-10)
    assert list(lo) == [0, 1, 1, 1, 4, 4, 4, 4, 4]
    vec1, chi1, _ = EP.epi_adapted(np.eye(9)[None])
    assert np.allclose(np.abs(vec1), np.eye(9)) and np.all(chi1 == 1.0)


def test_ring_projected_chain_equals_exact_off_site_average():
    """edc_green_proj (src/asr_edc.f90:174-232) with the dressing weights of spectral_dressing: on the ring (start site
    frozen to each species of its class, every target site and state) the projections of the exhausted chain give the
    conditional averages <[(z - H)^-1]_{0 iv, t l}> of the unscaled H = E_nu + h - h o h (brute force, 1e-7), the
    target = centre reproduces g; a planted mistake (the dressing without T_s) is detected."""
    from asr_amd import spectral as SP
    from asr_amd import workloads as W
    from exact import ring_graph, configurations, dense_s
    model, _ = ring_model()
    model = rotated(model, 1)       # site 0: binary class {1, 2}; sites 1 (ternary), 2 (binary), 3 (ordered)
    g = ring_graph(4)
    fl = np.ones(4, np.uint8)
    fl[0] = 0
    op = A.build_operator(g, model["cls"], model["nsc"], model["spc"], model["conc"], np.arange(4, dtype=np.int32)[None, :],
                          np.eye(9)[None], 0, 80, flippable=fl, block=model["block"], block0=model["block0"], nsp=4,
                          tsite=[0, 1, 2, 3])
    O = OracleOp(op)
    rng = np.random.default_rng(17)
    pp = dict(c=rng.uniform(-0.4, 0.4, (4, 3)), sqrdel=rng.uniform(0.5, 0.8, (4, 3)), enu=rng.uniform(-0.1, 0.1, (4, 3)),
              o=rng.uniform(-0.1, 0.05, (4, 3)))
    orb = {k: W.per_orbital(v) for k, v in pp.items()}
    s = dense_s(model)
    iv, z = 2, complex(0.1, 0.35)
    for cen in (1, 2):
        m = W.edc_member(pp, cen, iv, z)
        dc = orb["sqrdel"][cen] ** 2
        tw = np.stack([SP.spectral_dressing(dc, orb["sqrdel"] ** 2, op.spc[op.ncl[0] if t == 0 else model["cls"][t]],
                                            op.conc[model["cls"][t]], t == 0, op.nsmax) for t in range(4)])
        gg, gp, gp4, _ = O.edc_green_proj(m, z * m.pa + m.pb, iv, 1500, op.meta["tnode"], op.meta["tgop"], tw, z, tail="none")
        exact = np.zeros((4, 9), np.complex128)
        for sp, p in configurations(model, cen):
            c, d, o, e = (orb[k][sp].reshape(-1) for k in ("c", "sqrdel", "o", "enu"))
            h = d[:, None] * s * d[None, :] + np.diag(c)
            hm = np.diag(e) + h - h @ np.diag(o) @ h
            rhs = np.zeros(36, np.complex128)
            rhs[iv] = 1.0
            exact += p * np.linalg.solve(z * np.eye(36) - hm, rhs).reshape(4, 9)
        assert np.max(np.abs(gp - exact)) < 1e-7, (cen, np.max(np.abs(gp - exact)))
        assert abs(gp[0, iv] - gg) < 1e-9
        bad = tw.copy()
        bad[1:, :, :] = np.stack([SP.spectral_dressing(np.ones(9), np.ones((4, 9)), op.spc[model["cls"][t]],
                                                       op.conc[model["cls"][t]], False, op.nsmax) for t in range(1, 4)])
        _, gpb, _, _ = O.edc_green_proj(m, z * m.pa + m.pb, iv, 1500, op.meta["tnode"], op.meta["tgop"], bad, z, tail="none")
        assert np.max(np.abs(gpb - exact)) > 1e-4
