"""Parity at the measured sizes, not only at reduced ones.

* Headline (fcc ternary, 127,517 nodes, levels 4, D2h): all 54 chains x 30 levels of the GPU run against the oracle's
  chains, 1e-9 relative.
* The same chains on the reference's own group (D4h, orbital-permuting, 64,567 nodes): the 30 chains with a
  one-dimensional start orbital x 30 levels against the oracle, 1e-9 relative.
* Route ed3 on the headline operator: all 54 blocks x 2 chains x 30 levels in one batched run, three blocks (first,
  middle, last) against the oracle's ed3_blocks: a, b, c to 1e-9, the blocks to 1e-8.
* EPI pair blocks (fcc ternary first-shell pair, levels 3, 113,661 nodes): all 18 blocks x 2 chains x 30 in one batched
  run, three blocks against the oracle's epi_blocks: a, b, c to 1e-9, x to 1e-8.
* Projected chains on the headline operator (54 chains x 30, projections on the 19 sites of the start site's first two
  shells): three chains against the oracle's edc_green_proj: a, G, the projections to 1e-9.
* The L2_1 uneven batch (four sublattice operators with SRO classes, 256 chains of depth 20-60, one HIP graph): the
  deepest and the shallowest chain of every operator against the oracle at full depth, 1e-9.
* KPM on the 40^3-cell supercell (2,304,000 rows, 64 vectors, 2048 moments): size-independent properties, mu_0 = 1,
  |mu_n| <= 1, moment doubling = direct moments to 1e-12, the Jackson DOS integrates to 1; and the per-vector terms
  <r|t_n>/dim of the first two Rademacher vectors against the oracle's (16 moments), 1e-12.
* The adjoint (256 real chains x 40 on 127,517 nodes): gradient components of a contour objective against central finite
  differences of the GPU forward (h = 1e-5), 1e-6 relative.
"""
import numpy as np
import pytest

from asr_amd import workloads as W

pytestmark = pytest.mark.gpu


def _engine():
    from asr_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    return engine


def test_headline_all_chains_against_the_oracle():
    eng = _engine()
    from oracle.oracle import OracleOp
    wl = W.ternary_edc(L=8, levels=4, depth=30)
    print(f"headline: {wl.op.nn} nodes, {len(wl.members)} chains x 30", flush=True)
    G = eng.GpuOperator(wl.op)
    a, b, mu, _ = G.chains_edc(wl.members, wl.nlev)
    G.close()
    eps = np.stack([m.z * m.pa + m.pb for m in wl.members])
    ar, br, mr = OracleOp(wl.op).edc_batch(wl.members, eps, [m.l0 for m in wl.members], list(wl.nlev), project=True)
    assert np.array_equal(mu, mr)
    for j in range(len(wl.members)):
        n = int(mu[j])
        assert np.max(np.abs(a[j, :n] - ar[j, :n])) < 1e-9 * np.max(np.abs(ar[j, :n])), f"chain {j} a"
        assert np.max(np.abs(b[j, :n - 1] - br[j, :n - 1])) < 1e-9 * np.max(np.abs(br[j, :n - 1])), f"chain {j} b"


def test_d4h_headline_all_chains_against_the_oracle():
    eng = _engine()
    from oracle.oracle import OracleOp
    wl = W.ternary_edc(L=8, levels=4, depth=30, group="d4h")
    assert not np.all(wl.op.pi == np.arange(9)[None, :]), "D4h must permute orbitals"
    members = [m for m in wl.members if m.l0 in (0, 2, 4, 6, 8)]  # s, p_z, d_z2, d_xy, d_x2-y2
    nlev = np.full(len(members), 30, np.int32)
    print(f"d4h headline: {wl.op.nn} nodes, {len(members)} chains x 30", flush=True)
    G = eng.GpuOperator(wl.op)
    a, b, mu, _ = G.chains_edc(members, nlev)
    G.close()
    eps = np.stack([m.z * m.pa + m.pb for m in members])
    ar, br, mr = OracleOp(wl.op).edc_batch(members, eps, [m.l0 for m in members], list(nlev), project=True)
    assert np.array_equal(mu, mr)
    for j in range(len(members)):
        n = int(mu[j])
        assert np.max(np.abs(a[j, :n] - ar[j, :n])) < 1e-9 * np.max(np.abs(ar[j, :n])), f"chain {j} a"
        assert np.max(np.abs(b[j, :n - 1] - br[j, :n - 1])) < 1e-9 * np.max(np.abs(br[j, :n - 1])), f"chain {j} b"


def test_ed3_blocks_full_size_sample():
    eng = _engine()
    from oracle.oracle import OracleOp
    from asr_amd import ed3 as E3
    we = W.ternary_ed3(L=8, levels=4, depth=30)
    print(f"ed3: {we.op.nn} nodes, {len(we.members)} blocks x 2 chains x 30", flush=True)
    G = eng.GpuOperator(we.op)
    G.set_depth_tol(0.0, "hull")
    bk, a, b, c, mu, _ = G.ed3_blocks(we.members, we.info["cf"], we.info["oc"], 30)
    G.close()
    O = OracleOp(we.op)
    for j in (0, len(we.members) // 2, len(we.members) - 1):
        m = we.members[j]
        et, uu, qt = E3.ed3_coefs(we.info["ed"][j], m.z)
        bko, chains, _ = O.ed3_blocks(m, et, uu, qt, we.info["oc"][j], m.z, m.l0, 30, project=True, tail="hull")
        for sd in range(2):
            ao, bo, co, mo, closed = chains[sd]
            assert mu[j, sd] == mo == 30 and not closed
            assert np.max(np.abs(a[j, sd, :mo] - ao[:mo])) < 1e-9 * np.max(np.abs(ao[:mo])), (j, sd, "a")
            assert np.max(np.abs(b[j, sd, :mo - 1] - bo[:mo - 1])) < 1e-9 * np.max(np.abs(bo[:mo - 1])), (j, sd, "b")
            assert np.max(np.abs(c[j, sd, :mo] - co[:mo])) < 1e-9 * np.max(np.abs(co[:mo])), (j, sd, "c")
        assert np.max(np.abs(bk[j] - bko)) < 1e-8 * np.max(np.abs(bko)), (j, "blocks")


def test_epi_blocks_full_size_sample():
    eng = _engine()
    from oracle.oracle import OracleOp
    wp = W.fcc_epi_pair(L=8, levels=3)
    print(f"epi: {wp.op.nn} nodes, {len(wp.members)} blocks x 2 chains x 30", flush=True)
    G = eng.GpuOperator(wp.op)
    G.set_depth_tol(0.0, "hull")
    x, a, b, c, mu, _ = G.epi_blocks(wp.members, wp.info["v9"], 30)
    G.close()
    O = OracleOp(wp.op)
    for j in (0, len(wp.members) // 2, len(wp.members) - 1):
        m = wp.members[j]
        xo, chains = O.epi_blocks(m, m.z * m.pa + m.pb, wp.info["v9"][j], m.z, 30, project=True, tail="hull")
        for sd in range(2):
            ao, bo, co, mo, closed = chains[sd]
            assert mu[j, sd] == mo == 30 and not closed
            assert np.max(np.abs(a[j, sd, :mo] - ao[:mo])) < 1e-9 * np.max(np.abs(ao[:mo])), (j, sd, "a")
            assert np.max(np.abs(b[j, sd, :mo - 1] - bo[:mo - 1])) < 1e-9 * np.max(np.abs(bo[:mo - 1])), (j, sd, "b")
            assert np.max(np.abs(c[j, sd, :mo] - co[:mo])) < 1e-9 * np.max(np.abs(co[:mo])), (j, sd, "c")
        assert np.max(np.abs(x[j] - xo)) < 1e-8 * np.max(np.abs(xo)), (j, "x")


def test_projected_chains_full_size_sample():
    eng = _engine()
    from oracle.oracle import OracleOp
    ws = W.ternary_edc(L=8, levels=4, depth=30, targets=True)
    tn, tg, tw = ws.op.meta["tnode"], ws.op.meta["tgop"], ws.info["tw"]
    print(f"projected: {ws.op.nn} nodes, {len(ws.members)} chains x 30, {len(ws.info['tsite'])} targets", flush=True)
    G = eng.GpuOperator(ws.op)
    G.set_targets(tn, tg)
    G.set_depth_tol(0.0, "hull")
    g, gp, gp4, a, b, mu, _ = G.chains_edc_proj(ws.members, ws.nlev, tw)
    G.close()
    O = OracleOp(ws.op)
    for j in (0, len(ws.members) // 2, len(ws.members) - 1):
        m = ws.members[j]
        go, gpo, gp4o, (ao, bo, mo, _) = O.edc_green_proj(m, m.z * m.pa + m.pb, m.l0, int(ws.nlev[j]), tn, tg, tw[j], m.z,
                                                           tail="hull")
        assert mu[j] == mo
        assert np.max(np.abs(a[j, :mo] - ao[:mo])) < 1e-9 * np.max(np.abs(ao[:mo])), (j, "a")
        assert abs(g[j] - go) < 1e-9 * abs(go), (j, "G")
        assert np.max(np.abs(gp[j] - gpo)) < 1e-9 * np.max(np.abs(gpo)), (j, "projections")
        assert np.max(np.abs(gp4[j] - gp4o)) < 1e-9 * np.max(np.abs(gp4o)), (j, "projections, 4 levels back")


def test_l21_uneven_batch_full_depth_sample():
    eng = _engine()
    from oracle.oracle import OracleOp
    wls = W.l21_batch()
    groups = [(eng.GpuOperator(w.op), w.members, w.nlev) for w in wls]
    print(f"l21: {[(w.op.nn, len(w.members)) for w in wls]}", flush=True)
    res, _ = eng.chains_multi(groups, mode=2)
    for (G, members, nlev), w, (a, b, mu) in zip(groups, wls, res):
        O = OracleOp(w.op)
        for j in {int(np.argmax(nlev)), int(np.argmin(nlev))}:
            m = members[j]
            ea, eb, n = O.edc_chain(m, m.z * m.pa + m.pb, m.l0, int(nlev[j]))
            assert mu[j] == n == nlev[j]
            assert np.max(np.abs(a[j, :n] - ea[:n])) < 1e-9 * np.max(np.abs(ea[:n])), (w.name, j)
            assert np.max(np.abs(b[j, :n - 1] - eb[:n - 1])) < 1e-9 * np.max(np.abs(eb[:n - 1])), (w.name, j)
        G.close()


def test_kpm_full_size_properties():
    eng = _engine()
    from asr_amd import host as H
    g, species, par, block, block0 = W.kpm_supercell(L=40)
    R = eng.GpuRealization(g.nbr, species, par, block, block0)
    elo, ehi, _, _ = R.spectral_bounds(80, margin=0.0, nvec=1, seed=3)
    marg = 0.01 * (ehi - elo)
    c, s = 0.5 * (elo + ehi), 0.5 * (ehi - elo + 2.0 * marg) / 0.99
    print(f"kpm: {g.n * 9} rows, 64 vectors, 2048 moments", flush=True)
    mu = R.kpm_moments(c, s, 2048, 64, seed=5)["mu"]
    R.set_doubling(True)
    md = R.kpm_moments(c, s, 2048, 64, seed=5)["mu"]
    R.set_doubling(False)
    per = R.kpm_moments(c, s, 16, 64, seed=5, per_vector=True)["mu_per"][:2]
    R.close()
    from oracle.oracle import OracleReal
    _, per_o = OracleReal(g.nbr, species, par[0], par[1], par[2], par[3], block, block0).kpm_moments(
        c, s, 16, 2, 5, per_vector=True)
    assert np.max(np.abs(per - per_o)) < 1e-12, np.max(np.abs(per - per_o))
    assert abs(mu[0] - 1.0) < 1e-14
    assert np.max(np.abs(mu)) <= 1.0 + 1e-12
    assert np.max(np.abs(md - mu)) < 1e-12
    # the Jackson DOS over the scaled window integrates to g_0 mu_0 = 1 (Chebyshev-Gauss nodes: exact for the polynomial
    # part; the weight 1/sqrt(1 - x^2) is absorbed by the nodes)
    nq = 8192
    x = np.cos(np.pi * (np.arange(nq) + 0.5) / nq)
    rho = H.kpm_dos(mu, c, s, c + s * x)
    integral = float(np.sum(rho * np.sqrt(1.0 - x * x)) * np.pi / nq * s)
    assert abs(integral - 1.0) < 1e-10, integral


def test_adjoint_full_size_gradient_components():
    eng = _engine()
    import torch
    from asr_amd import adjoint as AD, host
    wa = W.ternary_adjoint()
    nlev = int(wa.nlev[0])
    nch = len(wa.members)
    G = eng.GpuOperator(wa.op)
    z, w = host.contour_closed(-1.6, -0.05, 0.6, 24, 8)
    zt, wt = torch.from_numpy(z), torch.from_numpy(w)
    weight = torch.from_numpy(np.asarray(wa.info["conc"])[[(j // 9) % 3 for j in range(nch)]] / 9.0)
    q0 = torch.full((nch,), 0.5, dtype=torch.float64)
    off = wa.info["c_offset"]

    def L(th):
        a, b = AD.chain_coefficients(th, G, wa.members, nlev, off)
        return AD.contour_objective(a, b, zt, wt, weight, q0, lam=3.0)

    print(f"adjoint: {wa.op.nn} nodes, {nch} chains x {nlev}", flush=True)
    theta = torch.tensor(wa.info["theta"], requires_grad=True)
    L(theta).backward()
    grad = theta.grad.numpy()
    h = 1e-5
    with torch.no_grad():
        for idx in [(0, 0, 0), (0, 2, 1), (1, 1, 2), (1, 0, 1)]:
            tp = wa.info["theta"].copy(); tp[idx] += h
            tm = wa.info["theta"].copy(); tm[idx] -= h
            fd = (float(L(torch.from_numpy(tp))) - float(L(torch.from_numpy(tm)))) / (2 * h)
            assert abs(grad[idx] - fd) < 1e-6 * max(abs(fd), np.max(np.abs(grad)) * 1e-2), (idx, grad[idx], fd)
    G.close()
