"""Golden values and closed forms of the reference that pin the host builder, the host terminators and the oracle
(no GPU).

* Published reduced node counts of the reference's example runs (example/reference/*/asr.out "nodes / sublattice").
* The screened fcc on-site block against the legacy INFO_FCC values (test/test_strconst.f90:10-33) and the analytic
  limits of the bare structure constants.
* The terminator tests of test/test_terminators.f90: the Nevanlinna bound on the two energy-dependent Cr e_g chains
  (hull obeys it, bp breaks it), the bp tail of cf_species against bp_asymptote, Beer-Pettifor on a uniform chain.
* KPM on a periodic sc band against the exact k-space spectrum of the same supercell (test/test_kpm.f90:17-43).
* The ordered nearest-neighbour band: a_0 = 0, b_0^2 = coordination (test_ordered_dos).
"""
import json
import math
import os

import numpy as np
import pytest

from asr_amd import augcls as A
from asr_amd import host
from asr_amd import structure as S

HERE = os.path.dirname(os.path.abspath(__file__))


# ------------------------------------------------------------------------------------------------ node counts
def _alloy_nodes(kind, L, nspecies, levels):
    vec, blk, r2max, w = S.lmto_alloy_blocks(kind, 1.2, S.STANDARD_ALPHA)
    g = S.periodic_graph_shells(kind, L, r2max)
    cls = np.zeros(g.n, np.int32)
    perm, mats = S.class_ops(g.coord, 2 * L, cls, "d4h")
    fl = np.ones(g.n, np.uint8)
    fl[0] = 0
    conc = [1.0 / nspecies] * nspecies
    op = A.build_operator(g, cls, [nspecies], [list(range(nspecies))], [conc], perm, S.orbital_rep(mats), 0, levels,
                          maxrank=1, flippable=fl)
    return op.nn


def test_published_node_count_fcc_ternary():
    """fe66ni14cr20_fcc: levels 3, rank 1, supercell 8, S^alpha kept to 1.2 alat -> 15827 (asr.out:113)."""
    assert _alloy_nodes("fcc", 8, 3, 3) == 15827


def test_published_node_count_bcc_binary():
    """cuzn_alloy: bcc A50B50, levels 4, rank 1, supercell 8 -> 17515 (asr.out:105)."""
    assert _alloy_nodes("bcc", 8, 2, 4) == 17515


@pytest.mark.parametrize("start", [0, 1])
def test_published_node_count_b2(start):
    """cuzn_b2: ordered B2 (one species per sublattice), levels 4, supercell 12 -> 197 on each sublattice (asr.out:111)."""
    vec, blk, r2max, w = S.lmto_alloy_blocks("bcc", 1.2, S.STANDARD_ALPHA)
    L = 12
    g = S.periodic_graph_shells("bcc", L, r2max)
    cls = ((g.coord[:, 0] + start) % 2).astype(np.int32)
    perm, mats = S.class_ops(g.coord, 2 * L, cls, "d4h")
    fl = np.ones(g.n, np.uint8)
    fl[0] = 0
    op = A.build_operator(g, cls, [1, 1], [[0], [1]], [[1.0], [1.0]], perm, S.orbital_rep(mats), 0, 4, maxrank=1,
                          flippable=fl)
    assert op.nn == 197


@pytest.mark.parametrize("wyckoff,shift,nodes", [("8c", 1, 8648), ("4a", 0, 4639), ("4b", 2, 4516)])
def test_published_node_count_l21(wyckoff, shift, nodes):
    """co2fe04cr06al_l21: Co on 8c, Al on 4a, Fe0.4Cr0.6 on 4b of the bcc-based L2_1 cell, levels 4, supercell 8:
    8648 / 4639 / 4516 nodes (asr.out:133).  Sublattice of a site: (x + y + z) mod 4 in half-lattice units."""
    vec, blk, r2max, w = S.lmto_alloy_blocks("bcc", 1.2, S.STANDARD_ALPHA)
    L = 8
    g = S.periodic_graph_shells("bcc", L, r2max)
    v = (g.coord.sum(axis=1) + shift) % 4
    cls = np.select([v == 0, (v == 1) | (v == 3), v == 2], [0, 1, 2]).astype(np.int32)
    perm, mats = S.class_ops(g.coord, 2 * L, cls, "d4h")
    fl = np.ones(g.n, np.uint8)
    fl[0] = 0
    op = A.build_operator(g, cls, [1, 1, 2], [[0], [1], [2, 3]], [[1.0], [1.0], [0.4, 0.6]], perm,
                          S.orbital_rep(mats), 0, 4, maxrank=1, flippable=fl)
    assert op.nn == nodes


# ------------------------------------------------------------------------------------------------ structure constants
def test_bare_structure_constant_limits():
    w = S.avg_ws_radius("fcc")
    d = np.array([0.5, 0.5, 0.0])
    b = S.bare_s0_block(d, w)
    assert abs(b[0, 0] - (-2.0 * w / math.sqrt(0.5))) < 1e-13
    bt = S.bare_s0_block(-d, w)
    assert np.max(np.abs(b - bt.T)) < 1e-12
    al = S.hardcore_alpha(0.7)
    assert abs(al[0] + al[1] + al[2] - (0.35 + 0.7 ** 3 / 6 + 0.7 ** 5 / 10)) < 1e-14


def test_screened_fcc_onsite_block_equals_legacy_info_fcc():
    """Hard-sphere screening sigma = 0.7 on the fcc cluster within 1.55 alat: the on-site diagonal equals the legacy
    INFO_FCC block (s 3.2326, p 3.0223, t2g 2.7422, eg 1.9492) to 2e-3 (test/test_strconst.f90:10-33)."""
    ref = np.array([3.2326, 3.0223, 3.0223, 3.0223, 2.7422, 2.7422, 1.9492, 2.7422, 1.9492])
    vec = S.lattice_cluster("fcc", 1.55)
    blk = S.screened_cluster(vec, S.avg_ws_radius("fcc"), S.hardcore_alpha(0.7))
    assert np.max(np.abs(np.diag(blk[0]) - ref)) < 2e-3


# ------------------------------------------------------------------------------------------------ terminators
@pytest.fixture(scope="module")
def cr_chains():
    with open(os.path.join(HERE, "golden", "cr_eg_chains.json")) as f:
        d = json.load(f)
    out = []
    for ch, zz in (("cr1", "zc1"), ("cr2", "zc2")):
        v = np.array(d[ch])
        out.append((v[:, 0] + 1j * v[:, 1], v[:, 2] + 1j * v[:, 3], complex(*d[zz])))
    return out


def _nevanlinna_ratio(cf, chains, tail):
    (a1, b1, z1), (a2, b2, z2) = chains
    g1 = cf(a1, b1, np.zeros(24), 24, z1, tail)[0]
    g2 = cf(a2, b2, np.zeros(24), 24, z2, tail)[0]
    return abs(g2 - g1) / abs(z2 - z1) / (min(-g1.imag, -g2.imag) / z1.imag)


def test_nevanlinna_bound_on_cr_eg_chains(cr_chains):
    """|dG/dz| <= -Im G / Im z holds for the hull tail (ratio ~0.4 in the reference) and fails for bp (~1000)."""
    from oracle.oracle import np_cf_species

    def host_cf(a, b, c, m, z, tail):
        return host.cf_species(a, b, c, m, z, tail)

    def orc_cf(a, b, c, m, z, tail):
        return np_cf_species(a, b, c, m, z, tail=tail)

    for cf in (host_cf, orc_cf):
        rh = _nevanlinna_ratio(cf, cr_chains, "hull")
        rb = _nevanlinna_ratio(cf, cr_chains, "bp")
        assert rh < 1.0, rh
        assert rb > 1.0, rb
    # host and oracle give the same terminated fractions on these chains
    for a, b, z in cr_chains:
        for tail in ("hull", "bp", "avg2", "none"):
            gh = host.cf_species(a, b, np.ones(24), 24, z, tail)
            go = np_cf_species(a, b, np.ones(24), 24, z, tail=tail)
            assert abs(gh[0] - go[0]) < 1e-10 * abs(go[0]) and abs(gh[1] - go[1]) < 1e-10 * abs(go[1]), tail


def test_cf_species_bp_tail_equals_bp_asymptote_on_real_chain():
    """test/test_terminators.f90:127-143: the bp tail of cf_species on a real chain equals the fraction closed with the
    real Beer-Pettifor asymptote (sqrt tail at (A, B)) to 1e-10."""
    a = np.array([0.10, -0.05, 0.20, 0.00, 0.15, 0.05])
    b = np.array([0.90, 1.10, 0.95, 1.05, 1.00, 1.00])
    ainf, binf, ok = host.bp_asymptote(a, b, 6)
    assert ok
    d = 0.0
    for ie in range(9):
        z = complex(-2.4 + 0.6 * ie, 0.05)
        g1, _ = host.cf_species(a, b, np.zeros(6), 6, z, "bp")
        dd = z - ainf
        s = np.sqrt(dd * dd - 4 * binf * binf)
        r1, r2 = (dd - s) / 2, (dd + s) / 2
        tail = r1 if abs(r1) <= abs(r2) else r2          # sqrt_tail: b^2 g of the decaying branch
        g2 = 1.0 / (z - a[5] - tail)
        for k in range(4, -1, -1):
            g2 = 1.0 / (z - a[k] - b[k] ** 2 * g2)
        d = max(d, abs(g1 - g2))
    assert d < 1e-10


def test_beer_pettifor_uniform_chain():
    a = np.zeros(60)
    b = np.ones(60)
    ainf, binf, _ = host.bp_asymptote(a, b, 56)
    assert abs(binf - 1.0) < 1e-2 and abs(ainf) < 1e-12


# ------------------------------------------------------------------------------------------------ KPM
def _sc_band(L):
    g = S.periodic_graph("sc", L)
    nslot = g.nbr.shape[1]
    block = np.zeros((nslot, 9, 9))
    block[:] = -np.eye(9)                     # every orbital an independent nearest-neighbour band, t = -1
    block0 = np.zeros((9, 9))
    par = np.zeros((4, 1, 9))
    par[1] = 1.0                              # h = c + d S d with c = 0, d = 1
    return g, np.zeros(g.n, np.uint8), par, block, block0


def _exact_sc_spectrum(nk, es, eta):
    k = 2 * np.pi * np.arange(nk) / nk
    eps = -2.0 * (np.cos(k)[:, None, None] + np.cos(k)[None, :, None] + np.cos(k)[None, None, :])
    eps = eps.ravel()
    return np.array([np.sum((eta / np.pi) / ((e - eps) ** 2 + eta ** 2)) for e in es]) / nk ** 3


def kpm_dos_error(mu):
    """Max relative error of the Lorentz-broadened (eta = 0.3) Jackson DOS against the exact spectrum (41 points)."""
    ne, ns = 1261, 41
    e = -6.3 + 12.6 * np.arange(ne) / (ne - 1)
    es = -6.0 + 12.0 * np.arange(ns) / (ns - 1)
    rho = host.kpm_dos(mu, 0.0, 6.3, e)
    de = e[1] - e[0]
    rhob = np.array([np.sum(rho * (0.3 / np.pi) / ((ei - e) ** 2 + 0.09) * de) for ei in e])
    ref = _exact_sc_spectrum(12, es, 0.3)
    idx = 1 + np.rint((es + 6.3) / 12.6 * (ne - 1)).astype(int) - 1
    return float(np.max(np.abs(rhob[idx] - ref)) / np.max(ref))


def test_kpm_oracle_reproduces_exact_sc_spectrum():
    """test/test_kpm.f90: 12^3 sc supercell, 400 moments, 60 random vectors, bounds [-6.3, 6.3]: within 2.5 %."""
    from oracle.oracle import OracleReal
    g, spec, par, block, block0 = _sc_band(12)
    O = OracleReal(g.nbr, spec, par[0], par[1], par[2], par[3], block, block0)
    mu = O.kpm_moments(0.0, 6.3, 400, 60, seed=5)
    assert abs(mu[0] - 1.0) < 1e-14
    assert kpm_dos_error(mu) < 2.5e-2


# ------------------------------------------------------------------------------------------------ ordered band
@pytest.mark.parametrize("kind,z", [("sc", 6), ("bcc", 8), ("fcc", 12)])
def test_ordered_nearest_neighbour_band(kind, z):
    """One species, hopping -1 between nearest neighbours on every orbital: a_0 = 0, b_0^2 = coordination."""
    from oracle.oracle import OracleOp
    L = 4
    g = S.periodic_graph(kind, L)
    cls = np.zeros(g.n, np.int32)
    perm, mats = S.class_ops(g.coord, 2 * L, cls, "identity")
    nslot = g.nbr.shape[1]
    block = -np.repeat(np.eye(9)[None], nslot, axis=0)
    op = A.build_operator(g, cls, [1], [[0]], [[1.0]], perm, S.orbital_rep(mats), 0, 3, maxrank=1,
                          block=block[None], block0=np.zeros((1, 9, 9)), nsp=1)
    m = A.Member(l0=0, c=np.zeros((1, 9)), d=np.ones((1, 9)), o=np.zeros((1, 9)), e=np.zeros((1, 9)))
    a, b, n = OracleOp(op).lanczos(m, 0, 3)
    assert abs(a[0]) < 1e-14
    assert abs(b[0] ** 2 - z) < 1e-12


def test_cf_closed_is_the_resolvent_of_the_chain_matrix():
    """cf_closed (src/asr_ed3.f90:214-230) of a complex-symmetric chain: g = [(z - J)^-1]_11 and f = sum_k c_k
    [(z - J)^-1]_k1 of the tridiagonal chain matrix J (dense solve), host C++ and the numpy oracle alike."""
    from oracle.oracle import np_cf_closed
    rng = np.random.default_rng(5)
    m = 11
    a = rng.normal(size=m) + 0.2j * rng.normal(size=m)
    b = 0.5 + rng.uniform(size=m) + 0.1j * rng.normal(size=m)
    c = rng.normal(size=m) + 1j * rng.normal(size=m)
    z = complex(0.3, 0.4)
    j = np.diag(a) + np.diag(b[:m - 1], 1) + np.diag(b[:m - 1], -1)
    col = np.linalg.solve(z * np.eye(m) - j, np.eye(m)[:, 0])
    g, f = host.cf_closed(a, b, c, m, z)
    assert abs(g - col[0]) < 1e-13 * abs(col[0])
    assert abs(f - c @ col) < 1e-12 * np.max(np.abs(c)) * np.max(np.abs(col))
    go, fo = np_cf_closed(a, b, c, m, z)
    assert abs(go - g) < 1e-14 and abs(fo - f) < 1e-13
    gs, fs = host.cf_species(a, b, c, m, z, "none")
    assert abs(gs - g) < 1e-14 and abs(fs - f) < 1e-13


def test_cf_species_batch_equals_single():
    """The batched host terminator (asr_host_cf_species_batch, the end of a contour pass) gives each chain's
    cf_species value bit for bit, for chains of different lengths and energies and every tail."""
    rng = np.random.default_rng(8)
    n, ld = 37, 24
    a = rng.uniform(-0.3, 0.3, (n, ld)) + 1j * rng.uniform(-0.05, 0.0, (n, ld))
    b = rng.uniform(0.2, 0.5, (n, ld)) + 1j * rng.uniform(-0.02, 0.02, (n, ld))
    m = rng.integers(1, ld + 1, n)
    z = rng.uniform(-0.6, 0.6, n) + 1j * rng.uniform(0.01, 0.3, n)
    for tail in ("hull", "avg2", "none"):
        g = host.cf_species_batch(a, b, m, z, tail)
        for j in range(n):
            gj, _ = host.cf_species(a[j], b[j], np.zeros(max(m[j], 1)), int(m[j]), z[j], tail)
            assert g[j] == gj, (tail, j)
