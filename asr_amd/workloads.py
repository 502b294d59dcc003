"""Seeded synthetic workloads of the chain engine (the model systems the throughput is quoted on).

    plain_binary      fcc A50B50, spd (9 orbitals), screened S^alpha to 2 shells (12 + 6 slots), c ~ U[-0.5, 0.5] Ry and
                      d ~ U[0.05, 0.15] per species and l, o = E_nu = 0 (H = h = c + d S d), one class, real chains of the
                      9 start orbitals (point group D2h: every orbital one-dimensional)
    ternary_edc       fcc Fe0.66Ni0.14Cr0.20-like ternary, 2 spin parameter sets, start site frozen to each species,
                      route-edc parameters (c, d, o, E_nu per species and spin mapped as edc_setup, src/asr_edc.f90:41-71),
                      9 orbitals x 3 species x 2 spins complex-symmetric chains at one contour energy
    l21_batch         4-sublattice L2_1-type bcc cell with partial disorder (1.0 / 0.5-0.5 / 0.4-0.6 / 0.3-0.3-0.4), one
                      operator per start sublattice, Warren-Cowley alpha_1 short-range order of the start site's first shell,
                      2 spins, chains of uneven depth
    ternary_adjoint   the ternary operator with real chains on H = c + d S d: 18 shared parameters (c, d per species and l),
                      each chain with its own fixed offset of c (its potential set), start site frozen to each species,
                      the 9 start orbitals; gradients of a contour objective through the chains
    kpm_supercell     fcc supercell of L^3 conventional cells (4 L^3 sites, 9 orbitals), seeded random ternary,
                      h = c + d S d with the same blocks
Species parameters are drawn per (species, l) and spread over the orbitals of each l (s | p x3 | d x5).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import structure as S
from .augcls import AugclsOperator, Member, build_operator, reorder_nodes

ORB_L = S.ORB_L


def per_orbital(v_l: np.ndarray) -> np.ndarray:
    """(nsp, 3) per-l values -> (nsp, 9) per orbital."""
    return np.ascontiguousarray(np.asarray(v_l)[:, ORB_L])


_BLOCKS = {}


def bond_blocks(kind: str, rcl: float = 1.0, alpha=S.STANDARD_ALPHA):
    key = (kind, rcl)
    if key not in _BLOCKS:
        vec, blk, r2max, w = S.lmto_alloy_blocks(kind, rcl, alpha)
        _BLOCKS[key] = (vec, blk, r2max)
    return _BLOCKS[key]


def graph_and_blocks(kind: str, L: int, rcl: float = 1.0):
    vec, blk, r2max = bond_blocks(kind, rcl)
    g = S.periodic_graph_shells(kind, L, r2max)
    block, block0 = S.bond_blocks_for(g.bvec, vec, blk)
    # symmetry zeros of the screened blocks are exact; rounding residue below 1e-14 is set to zero
    block[np.abs(block) < 1.0e-14] = 0.0
    block0 = block0.copy()
    block0[np.abs(block0) < 1.0e-14] = 0.0
    return g, block, block0


def draw_cd(rng: np.random.Generator, nsp: int):
    c = per_orbital(rng.uniform(-0.5, 0.5, (nsp, 3)))
    d = per_orbital(rng.uniform(0.05, 0.15, (nsp, 3)))
    return c, d


@dataclass
class Workload:
    name: str
    op: AugclsOperator
    members: list
    nlev: np.ndarray
    complex_route: bool
    info: dict


def plain_binary(L: int = 8, levels: int = 4, depth: int = 20, seed: int = 1, group: str = "d2h",
                 reorder: bool = True) -> Workload:
    rng = np.random.default_rng(seed)
    g, block, block0 = graph_and_blocks("fcc", L)
    cls = np.zeros(g.n, np.int32)
    perm, mats = S.class_ops(g.coord, 2 * L, cls, group)
    dmat = S.orbital_rep(mats)
    op = build_operator(g, cls, [2], [[0, 1]], [[0.5, 0.5]], perm, dmat, 0, levels, maxrank=1,
                        block=block[None], block0=block0[None], nsp=2)
    if reorder:
        op = reorder_nodes(op)
    c, d = draw_cd(rng, 2)
    z = np.zeros((2, 9))
    members = [Member(l0=l0, c=c, d=d, o=z, e=z) for l0 in range(9)]
    return Workload("fcc A50B50 plain ASR", op, members, np.full(9, depth, np.int32), False,
                    dict(L=L, levels=levels, group=group, seed=seed, nodes=op.nn))


def edc_params(rng: np.random.Generator, nsp: int, nspin: int):
    """Per spin: C - E_nu (c), sqrt(Delta) (d), E_nu and o per species and l (the gamma-to-alpha step is left out)."""
    out = []
    for _ in range(nspin):
        c = rng.uniform(-0.5, 0.5, (nsp, 3))
        sd = rng.uniform(0.05, 0.15, (nsp, 3))
        enu = rng.uniform(-0.1, 0.1, (nsp, 3))
        o = rng.uniform(-0.05, 0.05, (nsp, 3))
        out.append(dict(c=c, sqrdel=sd, enu=enu, o=o))
    return out


def edc_member(pp: dict, central: int, l0: int, z: complex) -> Member:
    """Species parameters for the central species as edc_setup (src/asr_edc.f90:41-71) and X(z) = z pa + pb (:253-258)."""
    nsp = pp["c"].shape[0]
    dc = per_orbital(pp["sqrdel"][central:central + 1] ** 2)[0]        # D_c (9)
    ds = per_orbital(pp["sqrdel"] ** 2)                                 # D_s (nsp, 9)
    c = dc[None, :] / ds * per_orbital(pp["c"])
    d = np.broadcast_to(np.abs(per_orbital(pp["sqrdel"][central:central + 1]))[0], (nsp, 9)).copy()
    o = per_orbital(pp["o"]) * ds / dc[None, :]
    e = np.zeros((nsp, 9))
    pa = 1.0 - dc[None, :] / ds
    pb = dc[None, :] * per_orbital(pp["enu"]) / ds
    return Member(l0=l0, c=c, d=d, o=o, e=e, frozen=central, pa=pa, pb=pb, z=z)


def ternary_edc(L: int = 8, levels: int = 4, depth: int = 30, seed: int = 2, group: str = "d2h",
                z: complex = complex(-0.2, 0.15), conc=(0.66, 0.14, 0.20), reorder: bool = True,
                targets: bool = False) -> Workload:
    """Route edc on the fcc ternary operator: 2 spins x 3 central species x 9 start orbitals.  targets: also the target
    nodes of the start site and its 18 neighbours (op.meta tnode / tgop) and per chain the dressing weights
    (info tw: (nchain, 19, nsmax, 9)) for the projected chains."""
    rng = np.random.default_rng(seed)
    g, block, block0 = graph_and_blocks("fcc", L)
    cls = np.zeros(g.n, np.int32)
    perm, mats = S.class_ops(g.coord, 2 * L, cls, group)
    dmat = S.orbital_rep(mats)
    fl = np.ones(g.n, np.uint8)
    fl[0] = 0                                    # the start site carries the central species (never flips)
    tsite = [0] + [int(v) for v in g.nbr[0]] if targets else None
    op = build_operator(g, cls, [3], [[0, 1, 2]], [list(conc)], perm, dmat, 0, levels, maxrank=1, flippable=fl,
                        block=block[None], block0=block0[None], nsp=3, tsite=tsite)
    if reorder:
        op = reorder_nodes(op)
    spins = edc_params(rng, 3, 2)
    members = [edc_member(spins[sp], cen, l0, z) for sp in range(2) for cen in range(3) for l0 in range(9)]
    info = dict(L=L, levels=levels, group=group, seed=seed, nodes=op.nn, z=z, params=spins)
    if targets:
        from . import spectral as SP
        tw = []
        for sp in range(2):
            dl = per_orbital(spins[sp]["sqrdel"] ** 2)
            for cen in range(3):
                dc = dl[cen]
                w = np.stack([SP.spectral_dressing(dc, dl, op.spc[0], op.conc[0], t == 0, op.nsmax) for t in range(len(tsite))])
                tw += [w] * 9
        info.update(tsite=np.asarray(tsite), tw=np.array(tw))
    return Workload("fcc ternary edc", op, members, np.full(len(members), depth, np.int32), True, info)


def ternary_ed3(L: int = 8, levels: int = 4, depth: int = 30, seed: int = 2, group: str = "d2h",
                z: complex = complex(-0.2, 0.15), conc=(0.66, 0.14, 0.20), reorder: bool = True,
                op: AugclsOperator | None = None) -> Workload:
    """Route ed3 on the fcc ternary operator: 2 spins x 3 central species x 9 start orbitals at one energy, each giving
    the two chains of ed3_blocks.  info: cf (nchain, 6, 3, 9) energy-linear coefficients, oc (nchain) o_c of the start
    orbital, ed (per chain the Ed3Species).  op: the operator of ternary_edc with the same L, levels, group and conc
    (its node tables do not depend on the species parameters), or None to build it."""
    from . import ed3 as E3
    rng = np.random.default_rng(seed)
    if op is None:
        g, block, block0 = graph_and_blocks("fcc", L)
        cls = np.zeros(g.n, np.int32)
        perm, mats = S.class_ops(g.coord, 2 * L, cls, group)
        dmat = S.orbital_rep(mats)
        fl = np.ones(g.n, np.uint8)
        fl[0] = 0
        op = build_operator(g, cls, [3], [[0, 1, 2]], [list(conc)], perm, dmat, 0, levels, maxrank=1, flippable=fl,
                            block=block[None], block0=block0[None], nsp=3)
        if reorder:
            op = reorder_nodes(op)
    spins = E3.ed3_params(rng, 3, 2)
    members, cf, oc, eds = [], [], [], []
    for sp in range(2):
        for cen in range(3):
            ed = E3.ed3_setup(spins[sp], cen)
            lin = E3.ed3_linear(ed)
            for l0 in range(9):
                members.append(E3.ed3_member(ed, l0, z))
                cf.append(lin)
                oc.append(E3.ed3_oc(ed, l0))
                eds.append(ed)
    return Workload("fcc ternary ed3", op, members, np.full(len(members), depth, np.int32), True,
                    dict(L=L, levels=levels, group=group, seed=seed, nodes=op.nn, z=z, cf=np.array(cf), oc=np.array(oc),
                         ed=eds))


def fcc_epi_pair(L: int = 8, levels: int = 4, seed: int = 31, z: complex = complex(-0.2, 0.15), lam: float = 0.5,
                 conc=(0.66, 0.14, 0.20), reorder: bool = True) -> Workload:
    """EPI blocks of a first-shell pair on the fcc ternary: site 0 on the path between species 0 and 1 at lam (species 3
    of the operator), the partner (first neighbour) frozen to species 0 or 2, the group the D2h operations that keep the
    partner (a class of its own), the nine symmetry-adapted start vectors (epi_adapted) per partner species.  info: v9
    (nchain, 9), par, pt, partner."""
    from . import epi as EP
    rng = np.random.default_rng(seed)
    g, block, block0 = graph_and_blocks("fcc", L)
    partner = int(g.nbr[0, 0])
    cls = np.zeros(g.n, np.int32)
    cls[partner] = 1
    perm, mats = S.class_ops(g.coord, 2 * L, cls, "d2h")
    dmat = S.orbital_rep(mats)
    fl = np.ones(g.n, np.uint8)
    fl[0] = 0
    fl[partner] = 0
    op = build_operator(g, cls, [3, 3], [[0, 1, 2], [0, 1, 2]], [list(conc), list(conc)], perm, dmat, 0, levels, maxrank=1,
                        flippable=fl, block=block[None], block0=block0[None], nsp=4, partner=partner)
    if reorder:
        op = reorder_nodes(op)
    par = [EP.epi_species(dict(enu=rng.uniform(-0.1, 0.1, 3), c=rng.uniform(-0.5, 0.5, 3), sqrdel=rng.uniform(0.05, 0.15, 3),
                               o=rng.uniform(-0.05, 0.05, 3))) for _ in range(3)]
    pt = EP.epi_path(par[0], par[1], lam)
    vec, chi, lo = EP.epi_adapted(dmat)
    members, v9 = [], []
    for sp in (0, 2):
        for a in range(9):
            members.append(EP.epi_member(par, pt, int(lo[a]), chi[:, a], sp, z))
            v9.append(vec[:, a])
    return Workload("fcc ternary EPI pair", op, members, np.full(len(members), 30, np.int32), True,
                    dict(L=L, levels=levels, seed=seed, nodes=op.nn, z=z, v9=np.array(v9), par=par, pt=pt, partner=partner,
                         nop=op.nop))


def ternary_adjoint(L: int = 8, levels: int = 4, nchains: int = 256, depth: int = 40, seed: int = 5, group: str = "d2h",
                    conc=(0.66, 0.14, 0.20), reorder: bool = True, offset: float = 0.02) -> Workload:
    """Real chains of the ternary operator for gradients: theta = (c, d) of shape (2, 3 species, 3 l shells).  Chain j
    starts at orbital j % 9 with the start site frozen to species (j // 9) % 3 and adds its own fixed offset to c."""
    rng = np.random.default_rng(seed)
    g, block, block0 = graph_and_blocks("fcc", L)
    cls = np.zeros(g.n, np.int32)
    perm, mats = S.class_ops(g.coord, 2 * L, cls, group)
    dmat = S.orbital_rep(mats)
    fl = np.ones(g.n, np.uint8)
    fl[0] = 0
    op = build_operator(g, cls, [3], [[0, 1, 2]], [list(conc)], perm, dmat, 0, levels, maxrank=1, flippable=fl,
                        block=block[None], block0=block0[None], nsp=3)
    if reorder:
        op = reorder_nodes(op)
    theta = np.stack([rng.uniform(-0.5, 0.5, (3, 3)), rng.uniform(0.05, 0.15, (3, 3))])
    dc = rng.uniform(-offset, offset, (nchains, 3, 3))
    members = []
    for j in range(nchains):
        c = per_orbital(theta[0] + dc[j])
        d = per_orbital(theta[1])
        z = np.zeros((3, 9))
        members.append(Member(l0=j % 9, c=c, d=d, o=z, e=z, frozen=(j // 9) % 3))
    return Workload("fcc ternary adjoint", op, members, np.full(nchains, depth, np.int32), False,
                    dict(L=L, levels=levels, group=group, seed=seed, nodes=op.nn, theta=theta, c_offset=dc,
                         conc=np.asarray(conc)))


def brick_order(g: S.Graph, brick: int = 4) -> S.Graph:
    """Renumber the sites of a periodic cubic supercell brick by brick (brick^3 conventional cells, bricks and cells in
    raster order): the neighbours of a brick then sit in a few nearby bricks, which keeps a sweep over the sites inside
    the L2 of one XCD.  The operator is the same up to the relabelling."""
    cell = g.coord // 2
    L = int(cell.max()) + 1
    bk = cell // brick
    nb = (L + brick - 1) // brick
    sub = (g.coord % 2) @ np.array([4, 2, 1])
    key = (((bk[:, 2] * nb + bk[:, 1]) * nb + bk[:, 0]) * brick ** 3 + ((cell[:, 2] % brick) * brick + cell[:, 1] % brick)
           * brick + cell[:, 0] % brick) * 8 + sub
    perm = np.argsort(key, kind="stable")            # new index -> old index
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    nbr = inv[g.nbr[perm]].astype(np.int32)
    return S.Graph(n=g.n, nbr=nbr, coord=g.coord[perm], bvec=g.bvec)


def kpm_supercell(L: int = 40, seed: int = 4, conc=(0.66, 0.14, 0.20), brick: int = 4):
    """Random ternary fcc supercell: (graph, species per site, par (4, nsp, 9), block, block0); sites in brick order."""
    rng = np.random.default_rng(seed)
    g, block, block0 = graph_and_blocks("fcc", L)
    if brick and L % brick == 0:
        g = brick_order(g, brick)
    species = rng.choice(len(conc), size=g.n, p=np.asarray(conc)).astype(np.uint8)
    c, d = draw_cd(rng, len(conc))
    par = np.zeros((4, len(conc), 9))
    par[0] = c
    par[1] = d
    return g, species, par, block, block0


def warren_cowley(rng: np.random.Generator, x, y, amp: float = 0.1) -> np.ndarray:
    """A seeded Warren-Cowley table alpha (n x m) with entries of order amp that satisfies both marginal conditions of two
    sublattices (sum_s y_s alpha_cs = 0 = sum_c x_c alpha_cs): a uniform draw projected onto them (the two projections
    commute), scaled down until every conditional concentration is non-negative."""
    x = np.asarray(x, dtype=float)
    y = np.asarray(y, dtype=float)
    a = rng.uniform(-amp, amp, (x.size, y.size))
    a = a - (a @ y)[:, None]
    a = a - (x @ a)[None, :]
    while np.any(y[None, :] * (1.0 - a) < 0.0):
        a *= 0.5
    return a


L21_SPC = [[0], [1, 2], [1, 3], [2, 3, 4]]
L21_CONC = [[1.0], [0.5, 0.5], [0.4, 0.6], [0.3, 0.3, 0.4]]


def l21_batch(L: int = 8, levels=(3, 3, 3, 2), nchains: int = 256, depth=(20, 60), seed: int = 3, sro: bool = True,
              group: str = "d2h", energies=(complex(-0.25, 0.12), complex(0.05, 0.12)), reorder: bool = True):
    """4-sublattice L2_1-type bcc cell (sublattice of a site: (x + y + z) mod 4 in half-lattice units) with occupations
    1.0 / 0.5-0.5 / 0.4-0.6 / 0.3-0.3-0.4 over 5 species, screened S^alpha to 2 bcc shells (8 + 6 slots).  One operator per
    start sublattice, the start site frozen to each of its species, 2 spins, 9 orbitals, 2 contour energies; with sro the 8
    first-shell sites of the start get shell classes whose concentrations are the Warren-Cowley conditionals of the central
    species (augcls_set_class_conc per chain, src/asr_class_scf.f90:1122-1126).  nchains of the 288 (sublattice, species,
    spin, orbital, energy) chains are drawn, depths uniform in [depth].  Returns (list of Workload per operator, info)."""
    rng = np.random.default_rng(seed)
    g, block, block0 = graph_and_blocks("bcc", L)
    base = (g.coord.sum(axis=1) % 4).astype(np.int32)
    nsp = 5
    spins = edc_params(rng, nsp, 2)
    shell = g.nbr[0, :8]
    # Warren-Cowley tables centre sublattice -> neighbour sublattice (alpha_1 in [-0.1, 0.1])
    wc = {(k1, k2): warren_cowley(rng, L21_CONC[k1], L21_CONC[k2]) for k1 in range(4) for k2 in range(4)}
    cands = []
    ops = []
    for kap in range(4):
        cls = ((base + kap) % 4).astype(np.int32)
        if sro:
            cls[shell] = 4 + cls[shell]
        nsc = [len(c) for c in L21_SPC] * (2 if sro else 1)
        spc = L21_SPC * (2 if sro else 1)
        conc = L21_CONC * (2 if sro else 1)
        perm, mats = S.class_ops(g.coord, 2 * L, cls, group)
        dmat = S.orbital_rep(mats)
        fl = np.ones(g.n, np.uint8)
        fl[0] = 0
        op = build_operator(g, cls, nsc, spc, conc, perm, dmat, 0, levels[kap], maxrank=1, flippable=fl,
                            block=block[None], block0=block0[None], nsp=nsp)
        if reorder:
            op = reorder_nodes(op)
        ops.append(op)
        for ic, cen in enumerate(L21_SPC[kap]):
            mconc = None
            if sro:
                mconc = [list(c) for c in conc]
                for lam in range(4):
                    cond, _ = S.sro_conditional(L21_CONC[kap], L21_CONC[lam], wc[(kap, lam)])
                    mconc[4 + lam] = list(cond[ic])
            for sp in range(2):
                for l0 in range(9):
                    for z in energies:
                        m = edc_member(spins[sp], cen, l0, z)
                        m.conc = mconc
                        cands.append((kap, m))
    pick = np.sort(rng.choice(len(cands), size=min(nchains, len(cands)), replace=False))
    depths = rng.integers(depth[0], depth[1] + 1, size=pick.size)
    wls = []
    for kap in range(4):
        idx = [q for q in range(pick.size) if cands[pick[q]][0] == kap]
        members = [cands[pick[q]][1] for q in idx]
        wls.append(Workload(f"L2_1 sublattice {kap}", ops[kap], members, depths[idx].astype(np.int32), True,
                            dict(L=L, levels=levels[kap], group=group, nodes=ops[kap].nn, sro=sro)))
    return wls
