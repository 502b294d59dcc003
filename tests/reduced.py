"""Models of the group-reduction checks: the same graph and classes built with the reference's point group (D4h about the
origin, `class_ops`, src/asr_symmetry.f90:67-92), with D2h and with the identity (the unreduced operator, every stabiliser
trivial, sq = 1), after the reference's own reduced-against-unreduced tests:

* ordered B2 on a bcc supercell of 4 (test/test_augcls_project.f90:26-73): 24 reduced nodes, 128 unreduced;
* the decorated bcc cell, L2_1 class pattern with the origin on each of three sublattices
  (test_slow/test_augcls.f90:208-290);
* the fcc ternary with a first-shell SRO class of conditional concentrations (test_slow/test_augcls.f90:470-519).

Species parameters are seeded draws in the ranges of the reference tests' potential parameters (edc_setup form, through
asr_amd.workloads.edc_member).
"""
from __future__ import annotations

import numpy as np

from asr_amd import augcls as A
from asr_amd import structure as S
from asr_amd import workloads as W

# the 1-D start orbitals of D4h in the reference's order s | y z x | xy yz 3z2-r2 xz x2-y2 (test_slow/test_augcls.f90:29
# takes 1, 3, 5, 7 one-based: s, z, xy, 3z2-r2)
L0S = [0, 2, 4, 6]
L0S_1D = [0, 2, 4, 6, 8]


def pot_params(nsp: int, seed: int) -> dict:
    """Per species and l: C - E_nu, sqrt(Delta), E_nu and o (the dictionary of workloads.edc_member)."""
    rng = np.random.default_rng(seed)
    return dict(c=rng.uniform(-0.3, 0.45, (nsp, 3)), sqrdel=rng.uniform(0.15, 0.3, (nsp, 3)),
                enu=rng.uniform(-0.1, 0.15, (nsp, 3)), o=rng.uniform(-0.05, 0.05, (nsp, 3)))


def _ops(g, cls, group, L):
    perm, mats = S.class_ops(g.coord, 2 * L, cls, group)
    return perm, S.orbital_rep(mats)


def b2_operator(group: str, levels: int = 6) -> A.AugclsOperator:
    """Ordered B2: class of a site = x mod 2 (one species per class), nothing flippable; blocks of build_lmto_alloy('bcc',
    4, rcl 1.2)."""
    L = 4
    g, block, block0 = W.graph_and_blocks("bcc", L, 1.2)
    cls = (g.coord[:, 0] % 2).astype(np.int32)
    perm, dmat = _ops(g, cls, group, L)
    fl = np.zeros(g.n, np.uint8)
    return A.build_operator(g, cls, [1, 1], [[0], [1]], [[1.0], [1.0]], perm, dmat, 0, levels, maxrank=1, flippable=fl,
                            block=block[None], block0=block0[None], nsp=2)


def l21_classes(coord: np.ndarray, sh) -> np.ndarray:
    """L2_1 pattern on the bcc graph (half-lattice units, period 4): all odd -> class 2; all even with (x + y + z) / 2 even
    -> class 0, odd -> class 1; coordinates shifted by sh (test_slow/test_augcls.f90:208-225)."""
    c = (coord + np.asarray(sh)[None, :]) % 4
    out = np.where(c[:, 0] % 2 == 1, 2, np.where((c.sum(axis=1) // 2) % 2 == 0, 0, 1))
    return out.astype(np.int32)


L21_SHIFTS = [(0, 0, 0), (2, 0, 0), (1, 1, 1)]


def decorated_bcc_operator(shift, group: str, levels: int = 3) -> A.AugclsOperator:
    """Classes 0 = ordered {0}, 1 = ternary {1, 2, 3} (0.3, 0.5, 0.2), 2 = ordered {2}; every site but the origin
    flippable; blocks of build_lmto_alloy('bcc', 4, rcl 1.05)."""
    L = 4
    g, block, block0 = W.graph_and_blocks("bcc", L, 1.05)
    cls = l21_classes(g.coord, shift)
    perm, dmat = _ops(g, cls, group, L)
    fl = np.ones(g.n, np.uint8)
    fl[0] = 0
    return A.build_operator(g, cls, [1, 3, 1], [[0], [1, 2, 3], [2]], [[1.0], [0.3, 0.5, 0.2], [1.0]], perm, dmat, 0,
                            levels, maxrank=1, flippable=fl, block=block[None], block0=block0[None], nsp=4)


FCC_X = np.array([0.3, 0.5, 0.2])


def fcc_sro_operator(group: str, shell: bool = True, levels: int = 3) -> A.AugclsOperator:
    """fcc ternary {1, 2, 3} (0.3, 0.5, 0.2) on one class; with shell, the 12 first-shell sites of the origin form a
    second class with the same species (their concentrations are set per chain, augcls_set_class_conc)."""
    L = 4
    g, block, block0 = W.graph_and_blocks("fcc", L, 1.05)
    cls = np.zeros(g.n, np.int32)
    if shell:
        r2 = np.sum(((g.coord + L) % (2 * L) - L) ** 2, axis=1)
        r2min = np.min(r2[1:])
        cls[1:][r2[1:] == r2min] = 1
        assert np.count_nonzero(cls) == 12
    perm, dmat = _ops(g, cls, group, L)
    fl = np.ones(g.n, np.uint8)
    fl[0] = 0
    nc = 2 if shell else 1
    return A.build_operator(g, cls, [3] * nc, [[1, 2, 3]] * nc, [list(FCC_X)] * nc, perm, dmat, 0, levels, maxrank=1,
                            flippable=fl, block=block[None], block0=block0[None], nsp=4)


def sro_conc(seed: int = 7):
    """Conditional first-shell concentrations cond[c, s] of a seeded Warren-Cowley table on the ternary (one
    sublattice: alpha symmetric, zero diagonal, the diagonal derived)."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(-0.15, 0.15, (3, 3))
    a = 0.5 * (a + a.T)
    np.fill_diagonal(a, 0.0)
    cond, _ = S.sro_conditional(FCC_X, FCC_X, a, same=True)
    return cond


def member(pp: dict, central: int, l0: int, z: complex, conc=None) -> A.Member:
    m = W.edc_member(pp, central, l0, z)
    if conc is not None:
        m.conc = conc
    return m


def ep_of(m: A.Member) -> np.ndarray:
    return m.z * m.pa + m.pb
