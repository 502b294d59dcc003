"""GPU parity of the SCF contour pass (asr_amd.scf.chain_greens / ed_moments; chain_greens / chain_worker / ed_moments,
src/asr_class_scf.f90:797-877, 966-1157) through the C ABI.

The ring of 4 sites with two class operators (start site in the binary class {1, 2}, start site in the ternary class
{0, 2, 3}) and four components, one with SRO concentrations: every chain task (contour energy, start orbital, component,
spin) of an exhausted chain (no tail) equals the brute-force configuration average of [(z - K)^-1]_{l0, l0} with the
component's parameters (edc: 1e-6), or of the blocks G_11, (hG)_11, (hGh)_11 of K = E O - H (ed3: 1e-6); the moments of the
GPU pass equal the moments of the brute-force G (1e-6).  The shares of three ranks sum to the one-rank arrays, each entry
non-zero in exactly one share (par_sum_z).
"""
import numpy as np
import pytest

from asr_amd import ed3 as E3
from asr_amd import scf
from asr_amd.workloads import edc_member

pytestmark = pytest.mark.gpu


def _engine():
    from asr_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    return engine


def _ring_classes():
    from exact import build_ring_operator, ring_model
    from test_oracle_exact import ring_ed3_params, rotated
    model, _ = ring_model()
    m1, m2 = rotated(model, 1), rotated(model, 2)
    ops = [build_ring_operator(m1, frozen_start=True), build_ring_operator(m2, frozen_start=True)]
    classes = [scf.ChainClass(ops[0], iorbk=[0, 2, 4, 7], worbk=[1, 3, 3, 2]), scf.ChainClass(ops[1])]
    # species 1, 2 start in the binary class (operator 0), species 0, 3 in the ternary class (operator 1); the last
    # component carries SRO concentrations for every class of its operator
    sro = [[1.0], [0.7, 0.3], [0.2, 0.2, 0.6]]
    comps = [scf.Component(1), scf.Component(0), scf.Component(0), scf.Component(1, conc=sro)]
    pp0 = ring_ed3_params()
    rng = np.random.default_rng(12)
    pp1 = {k: v + rng.uniform(-0.02, 0.02, v.shape) for k, v in pp0.items()}
    pp1["p"] = pp1["o"] ** 2 + np.abs(pp0["p"] - pp0["o"] ** 2)
    return model, [m1, m2], classes, comps, [pp0, pp1]


def _exact_edc(models, classes, comps, params, zs):
    from exact import exact_average_g
    nz = len(zs)
    gz = np.zeros((nz, 9, len(comps), len(params)), np.complex128)
    for gg, pp_ in enumerate(params):
        for pp, c in enumerate(comps):
            model = dict(models[c.pcl])
            if c.conc is not None:
                model["conc"] = c.conc
            for o, l0 in enumerate(classes[c.pcl].iorbk):
                m = edc_member(pp_, pp, l0, 0.0)
                par = dict(c=m.c, d=m.d, o=m.o, e=np.zeros_like(m.c), pa=m.pa, pb=m.pb)
                for j, z in enumerate(zs):
                    gz[j, o, pp, gg] = exact_average_g(model, par, z, l0, pp, ep=True)
    return gz


def test_chain_greens_edc_ring_exact_and_moments():
    """Route edc over a closed contour pass: gz of every task equals the brute-force average; the occupied moments, the
    DOS at E_F and the charges of the GPU pass equal those of the brute-force G; tasks of a start orbital the class
    does not have stay zero."""
    _engine()
    model, models, classes, comps, params = _ring_classes()
    enu = np.stack([np.stack([p["enu"][pp] for p in params]) for pp in range(4)])
    cp = scf.contour(0.05, enu, ebot_offset=1.0, ncont=16, ed_eta=0.05, fermi_kt=0.0)
    print(f"chain_greens edc ring: {cp.zq.size} energies, 2 spins, 4 components", flush=True)
    cg = scf.chain_greens(cp.zq, classes, comps, params, 1500, route="edc", depth_tol=0.0, tail="none")
    exact = _exact_edc(models, classes, comps, params, cp.zq)
    assert np.max(np.abs(cg.gz - exact)) < 1e-6, np.max(np.abs(cg.gz - exact))
    assert np.all(cg.gz[:, 4:, 1:3, :] == 0.0)
    assert cg.nchain == cp.zq.size * (9 + 4 + 4 + 9) * 2
    r = scf.ed_moments(cp, cg, classes, comps, enu)
    rx = scf.ed_moments(cp, scf.ChainGreens(exact, None, 0, 0, 0.0), classes, comps, enu)
    for f in ("mo", "dosv", "qc"):
        assert np.max(np.abs(getattr(r, f) - getattr(rx, f))) < 1e-6, f
    assert r.npsd == rx.npsd


def test_chain_greens_rank_shares_sum_to_the_full_pass():
    """Three ranks' shares (chain_worker's cyclic rule, computed one after the other on this GPU) hold disjoint non-zero
    entries and sum to the one-rank pass (par_sum_z semantics)."""
    _engine()
    model, models, classes, comps, params = _ring_classes()
    zs = np.array([complex(-0.3, 0.2), complex(0.0, 0.1), complex(0.2, 0.3)])
    full = scf.chain_greens(zs, classes, comps, params, 40, route="edc", depth_tol=0.0)
    parts = [scf.chain_greens(zs, classes, comps, params, 40, route="edc", depth_tol=0.0, rank=r, nrank=3).gz
             for r in range(3)]
    nz = np.array([p != 0 for p in parts]).sum(axis=0)
    assert np.all(nz == (full.gz != 0))
    # the batches differ between the shares (member tile, reduction order): rounding only
    assert np.max(np.abs(sum(parts) - full.gz)) < 1e-9 * np.max(np.abs(full.gz))


def test_chain_greens_ed3_ring_blocks_and_charges():
    """Route ed3: gb of every task equals the brute-force averages of G_11, (hG)_11, (hGh)_11; gz = bk0 + 2 o bk1 + p bk2 of
    the start orbital; the channel charges of ed_moments are the contour moments of gz with the orbital weights (the two
    combinations agree)."""
    _engine()
    from exact import exact_average_asa
    model, models, classes, comps, params = _ring_classes()
    comps = comps[:3]  # the blocks' brute force has no SRO variant
    enu = np.stack([np.stack([p["enu"][pp] for p in params]) for pp in range(3)])
    cp = scf.contour(0.05, enu, ebot_offset=1.0, ncont=6, ed_eta=0.05, fermi_kt=0.0, nvert=2)
    print(f"chain_greens ed3 ring: {cp.zq.size} energies", flush=True)
    # the ring's class operators know four species; the three components are the species 0..2 here
    cg = scf.chain_greens(cp.zq, classes, comps + [scf.Component(1)], params, 1500, route="ed3", depth_tol=0.0,
                          tail="none")
    for gg, pp_ in enumerate(params):
        orb = {k: E3.per_orbital(pp_[k]) for k in ("c", "enu", "o", "p")}
        orb["dl"] = E3.per_orbital(pp_["sqrdel"]) ** 2
        for pp, c in enumerate(comps):
            for o, l0 in enumerate(classes[c.pcl].iorbk):
                for j, z in enumerate(cp.zq):
                    ex = np.array(exact_average_asa(models[c.pcl], orb, z, l0, pp))
                    assert np.max(np.abs(cg.gb[j, o, pp, gg, :3] - ex)) < 1e-6, (gg, pp, l0, j)
                    assert abs(cg.gb[j, o, pp, gg, 3]) < 1e-6
    r = scf.ed_moments(cp, cg, classes, comps + [scf.Component(1)], np.concatenate([enu, enu[:1]]), params=params)
    for gg in range(2):
        for pp, c in enumerate(comps):
            cl = classes[c.pcl]
            for ll in range(3):
                q = sum(-wo * np.sum(cp.wq[:cp.nq] * cg.gz[:cp.nq, o, pp, gg]).imag / np.pi
                        for o, (wo, lo) in enumerate(zip(cl.worbk, cl.lorbk)) if lo == ll)
                assert abs(r.qc[pp, gg, ll] - q) < 1e-10 * max(1.0, abs(q)), (gg, pp, ll)


def test_window_run_on_the_ring_queue_on_device_equals_host():
    """Window runs of 16 slots over the ring's two class operators (24 and 16 nodes, one component with SRO
    concentrations), adaptive depth 1e-5: the pass with the queue on the device equals the pass with the queue followed
    by the host bit for bit, and both equal the same tasks in one batch each to the depth test's tolerance (a run in
    other slots rounds differently, and on this ring some tasks sit near a decision of the adaptive depth)."""
    eng = _engine()
    model, models, classes, comps, params = _ring_classes()
    zs = np.array([complex(-0.3 + 0.05 * k, 0.1 + 0.02 * k) for k in range(9)])
    ops = [eng.GpuOperator(c.op) for c in classes]
    kw = dict(route="edc", depth_tol=1e-5, tail="hull", gpu_ops=ops)
    ref = scf.chain_greens(zs, classes, comps, params, 60, max_batch=4096, window=False, **kw)
    dev = scf.chain_greens(zs, classes, comps, params, 60, max_batch=16, **kw)
    for G in ops:
        G.set_window_queue(False)
    host = scf.chain_greens(zs, classes, comps, params, 60, max_batch=16, **kw)
    for G in ops:
        G.set_window_queue(True)
        G.close()
    assert dev.nchain == ref.nchain and dev.nlev == ref.nlev
    assert np.array_equal(dev.gz, host.gz)
    assert np.max(np.abs(dev.gz - ref.gz)) < 1e-5 * np.max(np.abs(ref.gz))



def test_window_run_with_closed_chains_on_the_ring():
    """Chains that end early because b_k vanishes (the Krylov space is exhausted): on the ring with orbital 8 decoupled
    from the bonds (its rows and columns of the bond and on-site blocks zero) and the start site frozen, a chain from
    orbital 8 closes at its first level (h x = c x there); the others run their 40 levels.  A member that closes keeps a
    non-zero vector, so its slot is marked dirty and the next chain starts there from a full start column.  40 chains
    through 8 slots, the queue on the device and the host-followed queue: the depths equal the batched run's and the
    oracle's, and the coefficients agree with both to rounding.  The two queues do not agree bit for bit here: the host
    looks at the slots only at the levels where a member may leave by its depth or its test, so it hands a closed
    member's slot on later, and a chain that runs at other levels of the run rounds differently (lagged
    normalisation)."""
    eng = _engine()
    from exact import build_ring_operator, ring_model
    from oracle.oracle import OracleOp
    from test_oracle_exact import ring_ed3_params, rotated
    model, _ = ring_model()
    model = dict(model)
    blk, b0 = model["block"].copy(), model["block0"].copy()
    for B in (blk, b0):
        B[..., 8, :] = 0.0
        B[..., :, 8] = 0.0
    model["block"], model["block0"] = blk, b0
    op = build_ring_operator(rotated(model, 1), frozen_start=True)
    pp0 = ring_ed3_params()
    zs = [complex(-0.3 + 0.1 * k, 0.05 + 0.03 * k) for k in range(5)]
    # species 1 and 2 start in the binary class; orbital 8 on every third member
    members = [edc_member(pp0, 1 + (k // 3) % 2, 8 if k % 3 == 0 else k % 8, zs[k % 5]) for k in range(40)]
    G = eng.GpuOperator(op)
    G.set_tile(8)
    # a tolerance no chain meets: every member runs to its depth or closes
    G.set_depth_tol(1.0e-300, "none")
    a0, b0_, mu0, _ = G.chains_edc(members, 40, mode=0)
    G.set_window(8)
    res = {}
    for dev in (True, False):
        G.set_window_queue(dev)
        res[dev] = G.chains_edc(members, 40, mode=0)
    G.set_window_queue(True)
    G.set_window(0)
    G.set_depth_tol(0.0)
    G.close()
    print(f"ring depths {sorted(set(mu0.tolist()))}", flush=True)
    assert np.array_equal(mu0, np.where(np.arange(40) % 3 == 0, 1, 40))
    O = OracleOp(op)
    for a1, b1, mu1, _ in (res[True], res[False]):
        assert np.array_equal(mu1, mu0)
        for j in range(len(members)):
            k = int(mu0[j])
            sa = np.max(np.abs(a0[j, :k]))
            # 40 levels of a complex-symmetric chain: rounding grows to about 1e-11 of the coefficients' size
            assert np.max(np.abs(a1[j, :k] - a0[j, :k])) <= 1e-10 * sa, j
            if k > 1:
                sb = np.max(np.abs(b0_[j, :k - 1]))
                assert np.max(np.abs(b1[j, :k - 1] - b0_[j, :k - 1])) <= 1e-10 * sb, j
            if j % 5 == 0:
                m = members[j]
                ar, br, n = O.edc_chain(m, m.z * m.pa + m.pb, m.l0, 40)
                assert n == k, (j, n, k)
                assert np.max(np.abs(a1[j, :k] - ar[:k])) < 1e-9 * sa, j
                if k > 1:
                    assert np.max(np.abs(b1[j, :k - 1] - br[:k - 1])) < 1e-9 * sb, j
