"""Independent pins of the group-reduced operator and its projector (no GPU): the reduced chains against the unreduced
operator of the same graph, after the reference's own checks.

* test/test_augcls_project.f90:100-150: on the ordered B2 cell the 144-step edc chain of the D4h-reduced operator
  (24 nodes, 216 coordinates) equals the dense solve of z - H(z) on the unreduced operator (128 nodes, 1152 coordinates),
  1e-11; the projector is idempotent and symmetric, the operator keeps the symmetric subspace and is symmetric on it
  only (test_augcls_project.f90:74-96).
* test_slow/test_augcls.f90:238-290: decorated bcc (L2_1 class pattern, origin on each sublattice, D4h or D2d site
  group): reduced == unreduced chain coefficients and G, 1e-9.
* test_slow/test_augcls.f90:490-519: fcc ternary with a first-shell class of conditional SRO concentrations: reduced ==
  unreduced, 1e-9; shell concentrations = x reproduce the operator without a shell class, 1e-12.
* The headline's D2h reduction and the reference's D4h reduction of the fcc ternary against the unreduced node set.
* Planted mistakes in the reduced tables (an orbital sign sg, an orbit size sq, a stabiliser list) break the B2
  identity.

The unreduced operator is built with the identity group: every stabiliser is trivial, sq = 1, chi = sg = 1, so it carries
none of the tables under test; it is itself pinned by the ring brute force (test_oracle_exact.py).
"""
import dataclasses

import numpy as np
import pytest

from oracle.oracle import OracleOp, np_cf_species

import reduced as R


def g_none(a, b, m, z):
    """Continued fraction without a tail (ed_tail = 'none')."""
    return np_cf_species(a, b, np.zeros(max(m, 1)), m, z, "none")[0]


def dense_parts(O: OracleOp, mem):
    """Dense H (real) and the species-diagonal matrices of pa and pb of one member on an operator (columns = unit
    vectors through the oracle's apply and rand_c)."""
    n = O.op.dim
    h = np.zeros((n, n))
    xa = np.zeros((n, n), complex)
    xb = np.zeros((n, n), complex)
    pa = mem.pa.astype(complex)
    pb = mem.pb.astype(complex)
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        h[:, i] = O.apply(mem, e)
        xa[:, i] = O.rand_c(mem, pa, e.astype(complex))
        xb[:, i] = O.rand_c(mem, pb, e.astype(complex))
    return h, xa, xb


@pytest.fixture(scope="module")
def b2():
    opr = R.b2_operator("d4h")
    opu = R.b2_operator("identity")
    pp = R.pot_params(2, 20261110)
    return opr, opu, pp


def test_b2_node_sets(b2):
    opr, opu, _ = b2
    assert opr.nop == 16 and opr.nn == 24 and opr.dim == 216
    assert opu.nop == 1 and opu.nn == 128 and opu.dim == 1152
    assert np.all(np.diff(opu.stab_ptr) == 1) and np.all(opu.sq == 1.0)
    assert np.count_nonzero(np.diff(opr.stab_ptr) > 1) > 0


def test_b2_reduced_chain_equals_dense_unreduced_solve(b2):
    """edc chain of 144 steps on the D4h-reduced B2 operator == [z - H(z)]^-1 at the start component of the unreduced
    operator (test/test_augcls_project.f90:100-150), every one-dimensional start orbital, two energies."""
    opr, opu, pp = b2
    Or, Ou = OracleOp(opr), OracleOp(opu)
    worst = 0.0
    for l0 in R.L0S_1D:
        mu = R.member(pp, 0, l0, 0.0)
        h, xa, xb = dense_parts(Ou, mu)
        assert np.max(np.abs(h - h.T)) < 1e-13 * np.max(np.abs(h))
        for z in (complex(0.15, 0.03), complex(0.15, 0.003)):
            k = z * np.eye(opu.dim) - h - (z * xa + xb)
            rhs = np.zeros(opu.dim, complex)
            rhs[l0] = 1.0
            gex = np.linalg.solve(k, rhs)[l0]
            m = R.member(pp, 0, l0, z)
            a, b, mm = Or.edc_chain(m, R.ep_of(m), l0, 144, project=True)
            g = g_none(a, b, mm, z)
            worst = max(worst, abs(g - gex) / abs(gex))
    assert worst < 1e-11, worst


def _leak(O, m, x):
    return np.linalg.norm(x - O.project_c(m, x)) / np.linalg.norm(x)


def _hz(O, m, x):
    """H(z) x = H x + X(z) x for a complex x."""
    return O.apply(m, x.real) + 1j * O.apply(m, x.imag) + O.rand_c(m, R.ep_of(m).astype(complex), x)


def test_b2_projector_and_symmetric_subspace(b2):
    """(P) of test/test_augcls_project.f90:74-96 on the reduced B2 operator: P idempotent and symmetric in the bilinear
    form, a generic vector outside V, H(z) keeps V and is symmetric on V but not outside."""
    opr, _, pp = b2
    O = OracleOp(opr)
    rng = np.random.default_rng(3)
    for l0 in (0, 4, 8):
        m = R.member(pp, 0, l0, complex(0.15, 0.03))
        x = rng.uniform(-0.5, 0.5, opr.dim) + 1j * rng.uniform(-0.5, 0.5, opr.dim)
        y = rng.uniform(-0.5, 0.5, opr.dim) + 1j * rng.uniform(-0.5, 0.5, opr.dim)
        px, py = O.project_c(m, x), O.project_c(m, y)
        assert _leak(O, m, x) > 0.1
        assert _leak(O, m, px) < 1e-14
        assert abs(np.sum(y * px) - np.sum(py * x)) / abs(np.sum(py * x)) < 1e-13
        hx, hy = _hz(O, m, px), _hz(O, m, py)
        assert _leak(O, m, hx) < 1e-13
        assert abs(np.sum(py * hx) - np.sum(px * hy)) / abs(np.sum(px * hy)) < 1e-12
        gx, gy = _hz(O, m, x), _hz(O, m, y)
        assert abs(np.sum(y * gx) - np.sum(x * gy)) / abs(np.sum(x * gy)) > 1e-3


@pytest.mark.parametrize("shift", R.L21_SHIFTS)
def test_decorated_bcc_reduced_equals_unreduced(shift):
    """test_slow/test_augcls.f90:238-290: L2_1 classes with the origin on each sublattice; D4h on the rock-salt
    sublattices, D2d on the T_d one; every central species of the origin's class, the start orbitals s, z, xy, 3z2-r2,
    two energies, 8 levels: coefficients and G equal to 1e-9."""
    opr = R.decorated_bcc_operator(shift, "d4h")
    opu = R.decorated_bcc_operator(shift, "identity")
    ic = int(opr.ncl[0])
    assert opr.nop == (8 if ic == 2 else 16)
    assert opr.nn < opu.nn
    Or, Ou = OracleOp(opr), OracleOp(opu)
    pp = R.pot_params(4, 11)
    err = errg = 0.0
    for cen in opr.spc[ic, :opr.nsc[ic]]:
        for l0 in R.L0S:
            for kz in (1, 2):
                z = complex(-0.2 + 0.3 * kz, 0.15)
                m = R.member(pp, int(cen), l0, z)
                a1, b1, m1 = Ou.edc_chain(m, R.ep_of(m), l0, 8)
                a2, b2_, m2 = Or.edc_chain(m, R.ep_of(m), l0, 8)
                assert m1 == m2
                err = max(err, np.max(np.abs(a1[:m1] - a2[:m2])), np.max(np.abs(b1[:m1 - 1] - b2_[:m2 - 1])))
                errg = max(errg, abs(g_none(a1, b1, m1, z) - g_none(a2, b2_, m2, z)))
    assert err < 1e-9 and errg < 1e-9, (err, errg)


def test_fcc_sro_reduced_equals_unreduced():
    """test_slow/test_augcls.f90:490-519: the first-shell SRO class with the conditional concentrations of each central
    species: D4h-reduced == unreduced (1e-9); with the shell set to x the operator equals the one without a shell class
    (1e-12)."""
    opr = R.fcc_sro_operator("d4h")
    opu = R.fcc_sro_operator("identity")
    op0 = R.fcc_sro_operator("d4h", shell=False)
    assert opr.nop == op0.nop == 16
    Or, Ou, O0 = OracleOp(opr), OracleOp(opu), OracleOp(op0)
    cond = R.sro_conc()
    pp = R.pot_params(4, 12)
    x = list(R.FCC_X)
    err = err0 = 0.0
    for j, cen in enumerate((1, 2, 3)):
        for l0 in R.L0S:
            for kz in (1, 2):
                z = complex(-0.2 + 0.3 * kz, 0.15)
                m_x = R.member(pp, cen, l0, z, conc=[x, x])
                m_0 = R.member(pp, cen, l0, z)
                a2, b2_, m2 = Or.edc_chain(m_x, R.ep_of(m_x), l0, 8)
                a0, b0, m0 = O0.edc_chain(m_0, R.ep_of(m_0), l0, 8)
                err0 = max(err0, abs(g_none(a2, b2_, m2, z) - g_none(a0, b0, m0, z)))
                m_s = R.member(pp, cen, l0, z, conc=[x, list(cond[j])])
                a1, b1, m1 = Ou.edc_chain(m_s, R.ep_of(m_s), l0, 8)
                a2, b2_, m2 = Or.edc_chain(m_s, R.ep_of(m_s), l0, 8)
                err = max(err, np.max(np.abs(a1[:m1] - a2[:m2])), np.max(np.abs(b1[:m1 - 1] - b2_[:m2 - 1])),
                          abs(g_none(a1, b1, m1, z) - g_none(a2, b2_, m2, z)))
    assert err < 1e-9 and err0 < 1e-12, (err, err0)


@pytest.mark.parametrize("group", ["d2h", "d4h"])
def test_fcc_ternary_reduction_equals_unreduced(group):
    """The fcc ternary of the measured workloads (levels 2, frozen start site, every central species, the start
    orbitals that are one-dimensional under the group) reduced by D2h (the headline's group) and by D4h (the reference's)
    against the unreduced node set: real Lanczos and complex edc chains of 12 levels, 1e-9."""
    from asr_amd import workloads as W
    wr = W.ternary_edc(L=4, levels=2, depth=12, seed=12, group=group, reorder=False)
    wu = W.ternary_edc(L=4, levels=2, depth=12, seed=12, group="identity", reorder=False)
    assert wr.op.nn < wu.op.nn
    Or, Ou = OracleOp(wr.op), OracleOp(wu.op)
    l0s = list(range(9)) if group == "d2h" else R.L0S_1D
    err = 0.0
    for m in wr.members:
        if m.l0 not in l0s or (m.l0 + m.frozen) % 2:  # a sample of the 54 chains
            continue
        a1, b1, m1 = Ou.edc_chain(m, R.ep_of(m), m.l0, 12)
        a2, b2_, m2 = Or.edc_chain(m, R.ep_of(m), m.l0, 12)
        err = max(err, np.max(np.abs(a1 - a2)) / np.max(np.abs(a1)), np.max(np.abs(b1[:11] - b2_[:11])) / np.max(np.abs(b1)))
        r1, s1, _ = Ou.lanczos(m, m.l0, 12)
        r2, s2, _ = Or.lanczos(m, m.l0, 12)
        err = max(err, np.max(np.abs(r1 - r2)) / np.max(np.abs(r1)), np.max(np.abs(s1 - s2)) / np.max(np.abs(s1)))
    assert err < 1e-9, err


# ------------------------------------------------------------------------------------------------ planted mistakes
def _b2_error(opr, opu, pp, l0=4, z=complex(0.15, 0.03)):
    Or, Ou = OracleOp(opr), OracleOp(opu)
    mu = R.member(pp, 0, l0, z)
    h, xa, xb = dense_parts(Ou, mu)
    k = z * np.eye(opu.dim) - h - (z * xa + xb)
    rhs = np.zeros(opu.dim, complex)
    rhs[l0] = 1.0
    gex = np.linalg.solve(k, rhs)[l0]
    a, b, mm = Or.edc_chain(mu, R.ep_of(mu), l0, 144)
    return abs(g_none(a, b, mm, z) - gex) / abs(gex)


def _busiest_op(opr):
    hg = opr.hopg[opr.hopr >= 0]
    cnt = np.bincount(hg, minlength=opr.nop)
    cnt[0] = 0
    return int(np.argmax(cnt))


@pytest.mark.parametrize("mistake", ["sg", "sq", "stab"])
def test_planted_reduction_mistakes_break_the_b2_identity(b2, mistake):
    opr, opu, pp = b2
    assert _b2_error(opr, opu, pp) < 1e-11
    if mistake == "sg":
        # one orbital sign of the most used operation
        g = _busiest_op(opr)
        sg = opr.sg.copy()
        sg[g, 2] = -sg[g, 2]
        bad = dataclasses.replace(opr, sg=sg)
    elif mistake == "sq":
        # a wrong orbit size of the first node with a nontrivial orbit
        i = int(np.nonzero(opr.sq > 1.0)[0][0])
        sq = opr.sq.copy()
        sq[i] = np.sqrt(opr.sq[i] ** 2 / 2.0)
        bad = dataclasses.replace(opr, sq=sq)
    else:
        # the first nontrivial stabiliser list with its last operation replaced by one outside it (the projector of a
        # vector in V is exact for any subset of the stabiliser, so a dropped operation leaves the chain unchanged; a
        # wrong one does not)
        n = np.diff(opr.stab_ptr)
        i = int(np.nonzero((n > 1) & (n < opr.nop))[0][0])
        lst = opr.stab[opr.stab_ptr[i]:opr.stab_ptr[i + 1]]
        st = opr.stab.copy()
        st[opr.stab_ptr[i + 1] - 1] = [g for g in range(opr.nop) if g not in lst][0]
        bad = dataclasses.replace(opr, stab=st)
    assert _b2_error(bad, opu, pp) > 1e-6
