"""The SCF loop's contour pass on the chain engine (src/asr_class_scf.f90).

One pass of the self-consistency loop needs G(z) of every chain task (contour energy, start orbital, component, spin) and
the occupied moments of the channels from them:

    contour      the energies and weights of the pass (contour, src/asr_class_scf.f90:1414-1432): the closed arc + vertical
                 segment (contour_nodes_closed) or, with fermi_kt > 0, the Fermi-smeared line with its Matsubara poles
                 (contour_nodes_fermi), plus the E_F probe E_F + i ed_eta (weight 0) as node nq + 1;
    chain_greens gz(energy, start orbital, component, spin) = G(z) of the chain of each task, route edc (edc_green) or
                 route ed3 (gz = bk0 + 2 o bk1 + p bk2 of ed3_blocks, gb = the blocks) (chain_greens / chain_worker,
                 src/asr_class_scf.f90:966-1157);
    ed_moments   M_k = -(1/pi) Im sum_j w_j (z_j - E_nu)^k G(z_j) per component, spin and l channel with the start-orbital
                 weights, the DOS at E_F, the channel charges and the causality diagnostics (ed_moments :797-877).

The reference runs chain_worker as an OpenMP task loop over tasks t = 1 .. ntg * nomax * ncomp * nsp with a private class
operator per thread, re-parametrised, re-frozen and re-started per task.  Here every task of one class operator is a member
of a batched GPU call (asr_gpu_chains_edc / asr_gpu_ed3_blocks): the operator is resident once, the members carry their own
parameters, frozen species, class concentrations (SRO conditionals), start orbital and energy, and the batch is cut only by
a device-memory cap (the role of blk_nbmax).  With several ranks (rank, nrank) a rank computes the tasks the reference's
cyclic rule gives it and holds zeros elsewhere, so that one sum over the ranks (par_sum_z) gives every rank the full
arrays: each entry has exactly one non-zero term.

Species parameters are per spin dicts of (ncomp, 3) arrays per l in the screened (alpha) representation: c (= C - E_nu),
sqrdel, enu, o and, for route ed3, p; the species index is the component index (edc_setup / ed3_setup take pps(:, spin)
with the central species = the component).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import ed3 as E3
from . import host
from .augcls import AugclsOperator
from .workloads import edc_member

ORB_L = np.array([0, 1, 1, 1, 2, 2, 2, 2, 2])


@dataclass
class ChainClass:
    """One sublattice class operator and its start orbitals: iorbk (orbital index 0..8 of each start orbital), worbk (its
    weight: the number of symmetry-equivalent orbitals it stands for)."""
    op: AugclsOperator
    iorbk: list[int] = field(default_factory=lambda: list(range(9)))
    worbk: np.ndarray | None = None

    def __post_init__(self):
        self.iorbk = [int(v) for v in self.iorbk]
        self.worbk = np.ones(len(self.iorbk)) if self.worbk is None else np.asarray(self.worbk, float)
        if self.worbk.shape != (len(self.iorbk),):
            raise ValueError("ChainClass: one weight per start orbital")

    @property
    def lorbk(self) -> list[int]:
        return [int(ORB_L[o]) for o in self.iorbk]


@dataclass
class Component:
    """A (sublattice class, species) component: pcl (class index) and, with SRO, the class concentrations of the chains
    whose start site carries this component (augcls_set_class_conc with sro_cond, src/asr_class_scf.f90:1122-1126)."""
    pcl: int
    conc: list | None = None


@dataclass
class ContourPass:
    """Energies zq (nq + 1) and weights wq of one pass; node nq + 1 is the E_F probe (weight 0)."""
    zq: np.ndarray
    wq: np.ndarray
    nq: int


def contour(efv: float, enu, ebot_offset: float = 1.2, ncont: int = 32, ed_eta: float = 0.01, fermi_kt: float = 0.001,
            fermi_poles: int = 0, nvert: int = 8) -> ContourPass:
    """Nodes of one occupation pass (contour, src/asr_class_scf.f90:1414-1432): E_bot = min E_nu - ebot_offset; with
    fermi_kt > 0 the Fermi-smeared contour with npole Matsubara poles (fermi_npole unless fermi_poles > 0), ncont main-line
    nodes and 16 across the step; otherwise the closed arc (ncont) and vertical segment (nvert) at ed_eta."""
    ebot = float(np.min(enu)) - ebot_offset
    if fermi_kt > 0.0:
        npl = fermi_poles if fermi_poles > 0 else host.fermi_npole(fermi_kt)
        z, w = host.contour_fermi(ebot, efv, fermi_kt, npl, ncont, 16)
    else:
        z, w = host.contour_closed(ebot, efv, ed_eta, ncont, nvert)
    nq = z.size
    zq = np.append(z, complex(efv, ed_eta))
    wq = np.append(w, 0.0 + 0.0j)
    return ContourPass(zq, wq, nq)


def task_owner(t: int, ntg: int, nrank: int) -> int:
    """Rank of task t (0-based) under chain_worker's cyclic rule mod(t + t / ntg, nr) (src/asr_class_scf.f90:1110-1111,
    1-based there): the shift by the group index rotates the energies over the ranks."""
    return (t + t // ntg) % nrank


def _tasks(nz: int, classes: list[ChainClass], comps: list[Component], nsp: int, rank: int, nrank: int):
    """This rank's tasks (j, o, pp, gg) in chain_worker's order t = j + nz (o + nomax (pp + ncomp gg)) (one energy per
    task: the rule of eblock = 0)."""
    nomax = max(len(c.iorbk) for c in classes)
    ncomp = len(comps)
    out = []
    for t in range(nz * nomax * ncomp * nsp):
        j = t % nz
        o = (t // nz) % nomax
        pp = (t // (nomax * nz)) % ncomp
        gg = t // (nomax * nz * ncomp)
        if o >= len(classes[comps[pp].pcl].iorbk):
            continue  # this sublattice has fewer start orbitals
        if task_owner(t, nz, nrank) != rank:
            continue
        out.append((j, o, pp, gg))
    return out, nomax


def _member_bytes(op: AugclsOperator, ed3: bool) -> int:
    """Device bytes per chain member: the complex level vectors u, up, w, t1, t2, h1 (and tv for the blocks) of 9 (nn + 1)
    values, and the per-member partial sums."""
    return (7 if ed3 else 6) * (op.nn + 1) * 9 * 16 + 64 * 1024


@dataclass
class ChainGreens:
    """Output of chain_greens: gz (nz, nomax, ncomp, nsp); gb (nz, nomax, ncomp, nsp, 4) for route ed3 (else None);
    the chains and levels run (ed_nchain, ed_nlev) and the device time in ms."""
    gz: np.ndarray
    gb: np.ndarray | None
    nchain: int
    nlev: int
    device_ms: float


def chain_greens(zs, classes: list[ChainClass], comps: list[Component], params: list[dict], nlev: int,
                 route: str = "edc", depth_tol: float = 1.0e-5, tail: str = "hull", project: bool = True,
                 rank: int = 0, nrank: int = 1, mem_cap: float = 48.0 * 2 ** 30, max_batch: int = 112, mode: int = 0,
                 gpu_ops: list | None = None, window: bool = True) -> ChainGreens:
    """G(z) of every chain task at the energies zs (chain_greens / chain_worker, src/asr_class_scf.f90:966-1157).

    classes: the class operators (one per sublattice); comps: the components (their class and SRO concentrations);
    params: per spin the species parameters (ncomp, 3) per l; nlev: the chain depth; route 'edc' (edc_green, G from
    cf_species with the tail) or 'ed3' (ed3_blocks: gz = bk0 + 2 o_c bk1 + p_c bk2 of the start orbital, gb = bk);
    depth_tol / tail: the adaptive depth (ed_depth_tol, ed_tail); rank / nrank: this rank's share (zeros elsewhere);
    mem_cap: device bytes for the level vectors of one batched call, max_batch: members per call (0: the cap alone;
    112 by default: a node row of the member-fastest vectors stays within 16 KB, and batches of a few hundred members ran
    about 25 % slower per chain-level on the fcc ternary operator);
    mode: 0 batched, 2 one HIP graph per call; gpu_ops: resident GpuOperators of the classes (else made here and closed
    at the end; resident ones get their cached level vectors released).  Route edc with the adaptive depth tested on the
    device (tails hull, avg2, none) in mode 0 runs each class's tasks as one window run of max_batch member slots: a task
    takes the slot of one that has left, as chain_worker's dynamic schedule hands out the next task (window=False: batches).

    A class's tasks run energy-major (energy, spin, component, orbital), so that the members of one member tile share an
    energy: the adaptive depth retires them together and the tile leaves the level loop.  The results do not depend on
    the order (every chain is computed on its own)."""
    from .engine import GpuOperator
    if route not in ("edc", "ed3"):
        raise ValueError("chain_greens: route is 'edc' or 'ed3'")
    ed3 = route == "ed3"
    zs = np.asarray(zs, np.complex128).ravel()
    nz, ncomp, nsp = zs.size, len(comps), len(params)
    for pp, c in enumerate(comps):
        if not 0 <= c.pcl < len(classes):
            raise ValueError(f"chain_greens: component {pp} has no class {c.pcl}")
    for cl in classes:
        if cl.op.nsp != ncomp:
            raise ValueError("chain_greens: the species of every class operator are the components (nsp == ncomp)")
    tasks, nomax = _tasks(nz, classes, comps, nsp, rank, nrank)
    gz = np.zeros((nz, nomax, ncomp, nsp), np.complex128)
    gb = np.zeros((nz, nomax, ncomp, nsp, 4), np.complex128) if ed3 else None
    eds = {}
    if ed3:
        for gg in range(nsp):
            for pp in range(ncomp):
                eds[pp, gg] = E3.ed3_setup(params[gg], pp)
    own = gpu_ops is None
    gops = gpu_ops if gpu_ops is not None else [None] * len(classes)
    nchain = nlevs = 0
    dev_ms = 0.0
    try:
        for kp, cl in enumerate(classes):
            mine = sorted((t for t in tasks if comps[t[2]].pcl == kp), key=lambda t: (t[0], t[3], t[2], t[1]))
            if not mine:
                continue
            if gops[kp] is None:
                gops[kp] = GpuOperator(cl.op)
            G = gops[kp]
            G.set_depth_tol(depth_tol, tail)
            nbmax = max(1, int(mem_cap // _member_bytes(cl.op, ed3)))
            if max_batch > 0:
                nbmax = min(nbmax, max_batch)
            # route edc with the adaptive depth tested on the device, host-issued levels: one window run of nbmax slots
            # over the class's tasks (a chain takes the slot of one that has left); otherwise batches of nbmax
            win = window and (not ed3) and mode == 0 and depth_tol > 0.0 and tail != "bp" and len(mine) > nbmax
            G.set_window(nbmax if win else 0)
            step = len(mine) if win else nbmax
            if win:
                # the queue in order of the broadening |Im z|: the chains closest to the real axis run longest and go
                # first, those far from it (which leave after a few test levels) last, so that the run's tail is short
                mine.sort(key=lambda t: (abs(zs[t[0]].imag), t[0], t[3], t[2], t[1]))
            for q0 in range(0, len(mine), step):
                chunk = mine[q0:q0 + step]
                members, cf, oc = [], [], []
                for (j, o, pp, gg) in chunk:
                    l0 = cl.iorbk[o]
                    if ed3:
                        ed = eds[pp, gg]
                        m = E3.ed3_member(ed, l0, zs[j])
                        cf.append(E3.ed3_linear(ed))
                        oc.append(E3.ed3_oc(ed, l0))
                    else:
                        m = edc_member(params[gg], pp, l0, zs[j])
                    m.conc = comps[pp].conc
                    members.append(m)
                if ed3:
                    bk, _, _, _, mu, ms = G.ed3_blocks(members, np.array(cf), np.array(oc), nlev, project=project, mode=mode)
                    for q, (j, o, pp, gg) in enumerate(chunk):
                        ed = eds[pp, gg]
                        l0 = cl.iorbk[o]
                        gz[j, o, pp, gg] = bk[q, 0] + 2.0 * ed.o[pp, l0] * bk[q, 1] + ed.p[pp, l0] * bk[q, 2]
                        gb[j, o, pp, gg] = bk[q]
                    nlevs += int(mu.sum())
                else:
                    a, b, mu, ms = G.chains_edc(members, nlev, project=project, mode=mode)
                    g = host.cf_species_batch(a, b, mu, [zs[t[0]] for t in chunk], tail)
                    for q, (j, o, pp, gg) in enumerate(chunk):
                        gz[j, o, pp, gg] = g[q]
                    nlevs += int(mu.sum())
                nchain += len(chunk) * (2 if ed3 else 1)
                dev_ms += ms
            G.set_window(0)
    finally:
        for G in gops:
            if G is not None:
                if own:
                    G.close()
                else:
                    G.trim()  # the batches' cached level vectors
    return ChainGreens(gz, gb, nchain, nlevs, dev_ms)


@dataclass
class EdMoments:
    """Output of ed_moments: mo (ncomp, nsp, 3 [k], 3 [l]) occupied moments, dosv (ncomp, nsp, 3) DOS at E_F per channel,
    qc (ncomp, nsp, 3) channel charges, dm (ncomp, nsp, 3, 3) (route ed3: the DOS at E_F of each moment's combination),
    and the causality diagnostics: nnode nodes tested, npsd failures, asym_max, psd_max (route ed3)."""
    mo: np.ndarray
    dosv: np.ndarray
    qc: np.ndarray
    dm: np.ndarray
    nnode: int
    npsd: int
    asym_max: float
    psd_max: float


def ed_moments(cp: ContourPass, cg: ChainGreens, classes: list[ChainClass], comps: list[Component], enu,
               params: list[dict] | None = None) -> EdMoments:
    """Occupied moments of every (component, spin, l) channel from the G(z) of one pass (ed_moments,
    src/asr_class_scf.f90:797-877).  enu: (ncomp, nsp, 3) E_nu per l.  Route edc (cg.gb None):
        mo_k = -(1/pi) Im sum_o worbk_o sum_{j <= nq} w_j (z_j - E_nu,l)^k gz_j,  dosv = -(1/pi) sum_o worbk_o Im gz_{nq+1},
    qc = mo_0, and a node of the upper half plane with Im G >= 0 counts as a causality failure.  Route ed3: with the block
    moments xm_k and the E_F values xd_k of gb[..., k], o = o_alpha and p = p^alpha of the component (params):
        mo_0 += xm0 + 2 o xm1 + o^2 xm2,  mo_1 += xm1 + o xm2,  mo_2 += xm2  (dm the same with xd),
        qc += xm0 + 2 o xm1 + p xm2,  dosv += xd0 + 2 o xd1 + p xd2,
    and the 2x2 block -Im [[q0, q1 + o q2], [., q2]] (q0 = gb0 + 2 o gb1 + o^2 gb2) must be positive semidefinite at every
    node of the upper half plane; asym_max is the largest |q1a - q1b| / sqrt(|q0| |q2|)."""
    zq, wq, nq = cp.zq, cp.wq, cp.nq
    gz = cg.gz
    ncomp, nsp = len(comps), gz.shape[3]
    enu = np.asarray(enu, float).reshape(ncomp, nsp, 3)
    mo = np.zeros((ncomp, nsp, 3, 3))
    dosv = np.zeros((ncomp, nsp, 3))
    dm = np.zeros((ncomp, nsp, 3, 3))
    qc = np.zeros((ncomp, nsp, 3))
    up = zq.imag > 0.0
    nnode = npsd = 0
    asym_max = psd_max = 0.0
    if cg.gb is None:
        for gg in range(nsp):
            for pp in range(ncomp):
                cl = classes[comps[pp].pcl]
                for o, (wo, ll) in enumerate(zip(cl.worbk, cl.lorbk)):
                    g = gz[:, o, pp, gg]
                    nnode += int(np.count_nonzero(up))
                    npsd += int(np.count_nonzero(up & (g.imag >= 0.0)))
                    dosv[pp, gg, ll] -= wo * g[nq].imag / np.pi
                    for kk in range(3):
                        mo[pp, gg, kk, ll] -= wo * np.sum(wq[:nq] * (zq[:nq] - enu[pp, gg, ll]) ** kk * g[:nq]).imag / np.pi
        qc[...] = mo[:, :, 0, :]
        return EdMoments(mo, dosv, qc, dm, nnode, npsd, asym_max, psd_max)
    if params is None:
        raise ValueError("ed_moments: route ed3 needs the species parameters (o, p)")
    gb = cg.gb
    for gg in range(nsp):
        for pp in range(ncomp):
            oa = np.asarray(params[gg]["o"], float)[pp]
            pa = np.asarray(params[gg]["p"], float)[pp]
            cl = classes[comps[pp].pcl]
            for o, (wo, ll) in enumerate(zip(cl.worbk, cl.lorbk)):
                b = gb[:, o, pp, gg, :]
                xm = np.array([-wo * np.sum(wq[:nq] * b[:nq, k]).imag / np.pi for k in range(3)])
                xd = np.array([-wo * b[nq, k].imag / np.pi for k in range(3)])
                ol, pl = oa[ll], pa[ll]
                mo[pp, gg, 0, ll] += xm[0] + 2.0 * ol * xm[1] + ol ** 2 * xm[2]
                mo[pp, gg, 1, ll] += xm[1] + ol * xm[2]
                mo[pp, gg, 2, ll] += xm[2]
                dm[pp, gg, 0, ll] += xd[0] + 2.0 * ol * xd[1] + ol ** 2 * xd[2]
                dm[pp, gg, 1, ll] += xd[1] + ol * xd[2]
                dm[pp, gg, 2, ll] += xd[2]
                qc[pp, gg, ll] += xm[0] + 2.0 * ol * xm[1] + pl * xm[2]
                dosv[pp, gg, ll] += xd[0] + 2.0 * ol * xd[1] + pl * xd[2]
                for j in np.nonzero(up)[0]:
                    q0 = b[j, 0] + 2.0 * ol * b[j, 1] + ol ** 2 * b[j, 2]
                    a_, c_ = -q0.imag, -b[j, 2].imag
                    b_ = -(b[j, 1] + ol * b[j, 2]).imag
                    t_ = abs(b[j, 3]) / max(np.sqrt(abs(q0) * abs(b[j, 2])), np.finfo(float).tiny)
                    asym_max = max(asym_max, t_)
                    nnode += 1
                    if a_ <= 0.0 or c_ <= 0.0:
                        npsd += 1
                    else:
                        psd_max = max(psd_max, b_ ** 2 / (a_ * c_))
                        if b_ ** 2 > a_ * c_:
                            npsd += 1
    return EdMoments(mo, dosv, qc, dm, nnode, npsd, asym_max, psd_max)
