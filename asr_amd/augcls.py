"""The site-class augmented-space operator (the augcls_op of src/asr_augcls.f90:45-68) as flat arrays, and the per-chain
("member") parameters that the chain engine stacks.

An `AugclsOperator` holds everything that is shared by all chains on one node set: the node tables from the builder, the
orbital operation tables pi / sg (src/asr_augcls.f90:727-736), the bond blocks per site type and the class tables
(nsc, spc, mt).  A `Member` holds what differs between chains on the same node set: the start component, the characters
chi(g) of the start orbital (augcls_set_start, :107-118), the species parameters c, d, o, e (augcls_set_params, :121-126),
the frozen species of the start site (augcls_freeze_site, :129-141), the class concentrations (augcls_set_class_conc,
:178-185, for short-range order) and, for the energy-dependent route, the diagonal X(z) = z pa + pb (asr_edc.f90:253-258).

The local random weights wt(lam, l, k, class) = sum_s par(lam, spc(s)) mt(l, k, s, class) (local_weights, :362-377) are
computed here per member and parameter, so the kernels read one small table per member.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import host, structure


@dataclass
class AugclsOperator:
    nn: int
    nslot: int
    nop: int
    nsmax: int
    ncls: int
    nsp: int
    ntype: int
    hopr: np.ndarray      # (nn, nslot) int32, -1 = no link
    hopg: np.ndarray      # (nn, nslot) int32, operation mapping the neighbour node onto its representative
    flipr: np.ndarray     # (nn, nsmax) int32
    flipg: np.ndarray     # (nn, nsmax) int32
    lst: np.ndarray       # (nn,) local state of the node's own site
    ncl: np.ndarray       # (nn,) class of the node's site
    ntyp: np.ndarray      # (nn,) site type of the node's site (index into block / block0)
    site: np.ndarray      # (nn,) graph site
    level: np.ndarray     # (nn,)
    sq: np.ndarray        # (nn,) sqrt(orbit size)
    stab_ptr: np.ndarray  # (nn + 1,)
    stab: np.ndarray      # stabiliser operations
    pi: np.ndarray        # (nop, 9) int32: D(g)[pi(l), l] = sg(l)
    sg: np.ndarray        # (nop, 9) float64
    dmat: np.ndarray      # (nop, 9, 9)
    block: np.ndarray     # (ntype, nslot, 9, 9): block[t, k][l, l2] = S_{R, R + bvec(k)} for R of type t
    block0: np.ndarray    # (ntype, 9, 9)
    nsc: np.ndarray       # (ncls,)
    spc: np.ndarray       # (ncls, nsmax) global species (0-based), -1 beyond nsc
    conc: np.ndarray      # (ncls, nsmax)
    mt: np.ndarray        # (ncls, nsmax, nsmax, nsmax): mt[kap, l, k, s]
    start_site: int
    fz: np.ndarray        # (nn,) uint8: 1 node on the start site, 2 node on a frozen pair partner, 0 elsewhere
    meta: dict = field(default_factory=dict)

    @property
    def dim(self) -> int:
        return 9 * self.nn

    def chi_for(self, l0: int) -> np.ndarray:
        """Characters of the start orbital l0 (0-based): chi(g) = D(g)[l0, l0] = +-1 (augcls_set_start)."""
        d = self.dmat[:, l0, l0]
        if np.any(np.abs(np.abs(d) - 1.0) > 1.0e-9):
            raise ValueError("augcls_set_start: start orbital is not one-dimensional under the group")
        return np.sign(d).astype(np.float64)

    def with_blocks(self, block: np.ndarray, block0: np.ndarray) -> "AugclsOperator":
        self.block = np.ascontiguousarray(block, dtype=np.float64)
        self.block0 = np.ascontiguousarray(block0, dtype=np.float64)
        return self


def build_operator(graph: structure.Graph, cls: np.ndarray, nsc, spc, conc, perm: np.ndarray, dmat: np.ndarray, start: int,
                   maxlevel: int, maxrank: int = 0, flippable: np.ndarray | None = None, stype: np.ndarray | None = None,
                   block: np.ndarray | None = None, block0: np.ndarray | None = None, nsp: int | None = None,
                   nthreads: int = 0, partner: int | None = None, tsite=None) -> AugclsOperator:
    """Node set + tables (build_augcls_operator) and the orbital / class tables of the operator.  partner: a second
    non-flippable site whose nodes take a frozen species of their own (the partner of an EPI pair, augcls_freeze_any)."""
    if partner is not None and (flippable is None or flippable[partner] or partner == start):
        raise ValueError("build_augcls_operator: the frozen partner must be a non-flippable site other than the start")
    t = host.build_tables(graph.nbr, cls, perm, start, maxlevel, nsc, maxrank=maxrank, flippable=flippable, nthreads=nthreads,
                          tsite=tsite)
    nn = t["lst"].size
    nslot = graph.nbr.shape[1]
    nop = perm.shape[0]
    nsm = int(np.max(nsc))
    ncls = len(nsc)
    pi = np.zeros((nop, 9), np.int32)
    sg = np.zeros((nop, 9))
    for g in range(nop):  # D[pi(l), l] = sg(l): orbital operation tables (src/asr_augcls.f90:728-736)
        for l in range(9):
            hit = np.nonzero(np.abs(dmat[g, :, l]) > 0.5)[0]
            if hit.size == 0:
                raise ValueError("build_augcls_operator: orbital operation is not monomial")
            pi[g, l] = hit[-1]
            sg[g, l] = dmat[g, hit[-1], l]
    spc_a = -np.ones((ncls, nsm), np.int32)
    conc_a = np.zeros((ncls, nsm))
    mt = np.zeros((ncls, nsm, nsm, nsm))
    for kap in range(ncls):
        spc_a[kap, :nsc[kap]] = spc[kap][:nsc[kap]]
        conc_a[kap, :nsc[kap]] = conc[kap][:nsc[kap]]
        mt[kap] = structure.class_mt(conc_a[kap, :nsc[kap]], nsm) if nsc[kap] == nsm else _pad_mt(conc_a[kap, :nsc[kap]], nsm)
    if stype is None:
        stype = np.zeros(graph.n, np.int32)
    ntype = int(np.max(stype)) + 1
    if block is None:
        block = np.zeros((ntype, nslot, 9, 9))
    if block0 is None:
        block0 = np.zeros((ntype, 9, 9))
    if nsp is None:
        nsp = int(max(max(s[:n]) for s, n in zip(spc, nsc))) + 1
    site = t["site"]
    meta = {}
    if tsite is not None:
        # target nodes of the projected chains (tnode, tgop: (nt, nsmax), -1 = absent)
        meta.update(tsite=np.asarray(tsite, np.int32), tnode=t["tnode"], tgop=t["tgop"])
    return AugclsOperator(
        nn=nn, nslot=nslot, nop=nop, nsmax=nsm, ncls=ncls, nsp=nsp, ntype=ntype,
        hopr=t["hopr"], hopg=t["hopg"], flipr=t["flipr"], flipg=t["flipg"], lst=t["lst"],
        ncl=np.ascontiguousarray(cls[site], dtype=np.int32), ntyp=np.ascontiguousarray(stype[site], dtype=np.int32),
        site=site, level=t["level"], sq=np.sqrt(t["orb"].astype(np.float64)), stab_ptr=t["stab_ptr"], stab=t["stab"],
        pi=pi, sg=sg, dmat=np.ascontiguousarray(dmat, dtype=np.float64),
        block=np.ascontiguousarray(block, dtype=np.float64), block0=np.ascontiguousarray(block0, dtype=np.float64),
        nsc=np.asarray(nsc, np.int32), spc=spc_a, conc=conc_a, mt=mt, start_site=int(start),
        fz=(np.where(site == start, 1, 0) + (np.where(site == partner, 2, 0) if partner is not None else 0)).astype(np.uint8),
        meta=meta)


def reorder_nodes(op: AugclsOperator) -> AugclsOperator:
    """Relabel the nodes in reverse Cuthill-McKee order of the hop links, keeping the start node as node 0.

    The hop links (the S pass: 12-18 per node) keep the configuration of the environment, so they split the node set into
    sheets of one configuration class each; the flip links (the species-diagonal terms: 1-3 per node) join the sheets.
    Ordered by RCM of the hop links alone, about 90 % of the hop and flip partners of a run of 512 nodes lie inside that
    run (about 45 % in the builder's breadth-first order or in RCM of hops and flips together), so the gathers of one
    sweep stay in the L2 of one XCD.  The operator is the same up to the relabelling (chain coefficients agree to
    rounding)."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    nn = op.nn
    rows = np.repeat(np.arange(nn), op.nslot)
    cols = op.hopr.reshape(-1)
    ok = cols >= 0
    a = sp.csr_matrix((np.ones(int(ok.sum())), (rows[ok], cols[ok])), shape=(nn, nn))
    a = (a + a.T).tocsr()
    p = np.asarray(reverse_cuthill_mckee(a, symmetric_mode=True), dtype=np.int64)
    p = np.concatenate([[0], p[p != 0]])           # new index -> old index, the start node first
    inv = np.empty(nn, np.int64)
    inv[p] = np.arange(nn)

    def relink(t):
        t = t[p]
        return np.where(t >= 0, inv[np.maximum(t, 0)], -1).astype(np.int32)

    counts = np.diff(op.stab_ptr)[p]
    stab_ptr = np.zeros(nn + 1, np.int32)
    stab_ptr[1:] = np.cumsum(counts)
    stab = np.concatenate([op.stab[op.stab_ptr[o]:op.stab_ptr[o + 1]] for o in p]).astype(np.int32) if op.stab.size else \
        op.stab.copy()
    import dataclasses
    meta = dict(op.meta)
    meta["node_order"] = "rcm of the hop links"
    if "tnode" in meta:
        tn = meta["tnode"]
        meta["tnode"] = np.where(tn >= 0, inv[np.maximum(tn, 0)], -1).astype(np.int32)
    return dataclasses.replace(op, hopr=relink(op.hopr), hopg=np.ascontiguousarray(op.hopg[p]), flipr=relink(op.flipr),
                               flipg=np.ascontiguousarray(op.flipg[p]), lst=np.ascontiguousarray(op.lst[p]),
                               ncl=np.ascontiguousarray(op.ncl[p]), ntyp=np.ascontiguousarray(op.ntyp[p]),
                               site=np.ascontiguousarray(op.site[p]), level=np.ascontiguousarray(op.level[p]),
                               sq=np.ascontiguousarray(op.sq[p]), stab_ptr=stab_ptr, stab=stab,
                               fz=np.ascontiguousarray(op.fz[p]), meta=meta)


def _pad_mt(conc, nsm):
    mt = np.zeros((nsm, nsm, nsm))
    n = conc.size
    mt[:n, :n, :n] = structure.class_mt(conc, n)
    return mt


def local_weights(op: AugclsOperator, par: np.ndarray, mt: np.ndarray | None = None) -> np.ndarray:
    """wt[kap, l, k, lam] = sum_s par[spc(s, kap), lam] mt[kap, l, k, s] (local_weights, src/asr_augcls.f90:362-377);
    par (nsp, 9)."""
    mt = op.mt if mt is None else mt
    wt = np.zeros((op.ncls, op.nsmax, op.nsmax, 9))
    for kap in range(op.ncls):
        for kk in range(op.nsc[kap]):
            for k in range(op.nsc[kap]):
                for s in range(op.nsc[kap]):
                    wt[kap, k, kk, :] += par[op.spc[kap, s], :] * mt[kap, k, kk, s]
    return wt


@dataclass
class Member:
    """One chain on an operator: start component iv = 9 * node + l0 (node 0 = start node), species parameters (nsp, 9),
    frozen species of the start site (-1: none), per-class concentrations (None: the operator's)."""
    l0: int
    c: np.ndarray
    d: np.ndarray
    o: np.ndarray
    e: np.ndarray
    frozen: int = -1
    conc: np.ndarray | None = None
    pa: np.ndarray | None = None   # energy-dependent diagonal X(z) = z pa + pb (nsp, 9), route edc
    pb: np.ndarray | None = None
    z: complex = 0.0
    depth: int = 0
    frozen2: int = -1              # species of the frozen pair partner (fz = 2 nodes), -1: none
    chi: np.ndarray | None = None  # characters of a symmetry-adapted start vector (nop,), None: those of orbital l0

    def mt(self, op: AugclsOperator) -> np.ndarray:
        if self.conc is None:
            return op.mt
        mt = np.zeros_like(op.mt)
        for kap in range(op.ncls):
            mt[kap] = _pad_mt(np.asarray(self.conc[kap][:op.nsc[kap]], float), op.nsmax)
        return mt


def member_tables(op: AugclsOperator, members: list[Member]) -> dict:
    """Per-member arrays in the layout of the chain engine: chi (nb, nop), weights (nb, 5, ncls, nsmax, nsmax, 9) for
    (c, d, o, e, pa) and pb separately, frozen parameter rows (nb, 6, 9), frozen flags, start components."""
    nb = len(members)
    chi = np.zeros((nb, op.nop))
    wts = np.zeros((nb, 6, op.ncls, op.nsmax, op.nsmax, 9))
    fpar = np.zeros((nb, 6, 9))
    frozen = np.zeros(nb, np.int32)
    iv = np.zeros(nb, np.int32)
    uniform_d = np.ones(nb, np.int32)
    zero_o = np.ones(nb, np.int32)
    for j, m in enumerate(members):
        chi[j] = op.chi_for(m.l0)
        mt = m.mt(op)
        pars = [m.c, m.d, m.o, m.e,
                m.pa if m.pa is not None else np.zeros((op.nsp, 9)),
                m.pb if m.pb is not None else np.zeros((op.nsp, 9))]
        for q, par in enumerate(pars):
            par = np.asarray(par, dtype=np.float64).reshape(op.nsp, 9)
            wts[j, q] = local_weights(op, par, mt)
            if m.frozen >= 0:
                fpar[j, q] = par[m.frozen]
        frozen[j] = 1 if m.frozen >= 0 else 0
        iv[j] = m.l0
        uniform_d[j] = int(np.all(np.asarray(m.d).reshape(op.nsp, 9) == np.asarray(m.d).reshape(op.nsp, 9)[0:1]))
        zero_o[j] = int(not np.any(np.asarray(m.o) != 0.0))
    return dict(chi=chi, wts=wts, fpar=fpar, frozen=frozen, iv=iv, uniform_d=uniform_d, zero_o=zero_o)
