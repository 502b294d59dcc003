"""GPU parity of the EPI pair blocks (asr_gpu_epi_blocks; epi_blocks / epi_blocks_blk, src/asr_epi.f90:628-759) through
the C ABI: two frozen sites (site 0 on the path point, the pair partner frozen to a species), symmetry-adapted start
combinations with their characters under the pair group.

Bar: both chains' a_n, b_n, c_n to 1e-9 relative against the oracle, the blocks x to 1e-8; on the ring the blocks equal the
brute-force conditional averages of v9^T G_00 v9, v9^T (hG)_00 v9, v9^T (hGh)_00 v9 (1e-6, exhausted chains, no tail).
"""
import numpy as np
import pytest

from asr_amd import epi as EP
from asr_amd import workloads as W

pytestmark = pytest.mark.gpu


def _engine():
    from asr_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    return engine


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300))


@pytest.fixture(scope="module")
def epi_small():
    return W.fcc_epi_pair(L=4, levels=2, seed=31)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_epi_blocks_match_oracle(epi_small, mode):
    """18 blocks (nine adapted start vectors x two partner species) on the fcc pair operator in every mode against the
    oracle's epi_blocks: coefficients, projections, stopping levels and x."""
    eng = _engine()
    from oracle.oracle import OracleOp
    wl = epi_small
    nlev = 14
    G = eng.GpuOperator(wl.op)
    G.set_depth_tol(0.0, "hull")
    print(f"epi mode {mode}: {len(wl.members)} blocks, {wl.op.nn} nodes, {wl.op.nop} operations", flush=True)
    x, a, b, c, mu, _ = G.epi_blocks(wl.members, wl.info["v9"], nlev, mode=mode)
    O = OracleOp(wl.op)
    for j, m in enumerate(wl.members):
        xo, chains = O.epi_blocks(m, m.z * m.pa + m.pb, wl.info["v9"][j], m.z, nlev, project=True, tail="hull")
        for sd in range(2):
            ao, bo, co, mo, closed = chains[sd]
            assert mu[j, sd] == mo
            assert rel(a[j, sd, :mo], ao[:mo]) < 1e-9, (j, sd)
            assert rel(b[j, sd, :mo - 1], bo[:mo - 1]) < 1e-9, (j, sd)
            assert rel(c[j, sd, :mo], co[:mo]) < 1e-9, (j, sd)
        assert rel(x[j], xo) < 1e-8, (j, x[j], xo)


def test_epi_ring_exact_average_on_gpu():
    """The ring with two frozen sites through the GPU: the blocks of exhausted chains are the brute-force conditional
    averages; a wrong partner species is detected."""
    eng = _engine()
    from exact import exact_average_pair_blocks
    from test_oracle_exact import _ring_pair_setup
    model, op, par, pt, sites, partner = _ring_pair_setup()
    rng = np.random.default_rng(5)
    v9s = [np.eye(9)[4], rng.standard_normal(9), rng.standard_normal(9)]
    v9s = [v / np.linalg.norm(v) for v in v9s]
    members, cases = [], []
    for k, v9 in enumerate(v9s):
        z = complex(-0.25 + 0.2 * k, 0.3)
        members.append(EP.epi_member(par, pt, 4, np.ones(1), 2, z))
        cases.append((v9, z))
    G = eng.GpuOperator(op)
    G.set_depth_tol(0.0, "none")
    print(f"epi ring: {len(members)} blocks, {op.nn} nodes, 1500 levels", flush=True)
    x, a, b, c, mu, _ = G.epi_blocks(members, np.array(v9s), 1500, mode=2)
    for j, (v9, z) in enumerate(cases):
        exact = exact_average_pair_blocks(model, sites, z, v9, partner, 2, 4)
        assert np.max(np.abs(x[j, :3] - exact)) < 1e-6, (j, x[j], exact)
        assert abs(x[j, 3]) < 1e-6
    wrong = [EP.epi_member(par, pt, 4, np.ones(1), 1, cases[0][1])]
    xw, *_ = G.epi_blocks(wrong, v9s[0][None], 1500, mode=0)
    exact = exact_average_pair_blocks(model, sites, cases[0][1], v9s[0], partner, 2, 4)
    assert np.max(np.abs(xw[0, :3] - exact)) > 1e-4


def test_epi_integrand_closes_on_the_blocks(epi_small):
    """The trace of epi_integrand over the nine adapted vectors from the GPU blocks equals the same trace from the
    oracle's blocks (the host formula on both), for both partner species."""
    eng = _engine()
    from oracle.oracle import OracleOp
    wl = epi_small
    G = eng.GpuOperator(wl.op)
    x, *_ = G.epi_blocks(wl.members, wl.info["v9"], 12)
    O = OracleOp(wl.op)
    pt = wl.info["pt"]
    for half in range(2):
        tg = to = 0.0
        for a in range(9):
            j = 9 * half + a
            m = wl.members[j]
            xo, _ = O.epi_blocks(m, m.z * m.pa + m.pb, wl.info["v9"][j], m.z, 12)
            tg += EP.epi_integrand(pt, m.l0, m.z, x[j])
            to += EP.epi_integrand(pt, m.l0, m.z, xo)
        assert abs(tg - to) < 1e-8 * abs(to)


def test_epi_thermodynamic_integration_on_ring():
    """The EPI driver on the ring (src/asr_epi.f90:4-24): the difference of dOmega/dlambda between two partner species
    (epi_domega over a Fermi contour, exhausted chains) equals the brute-force Hellmann-Feynman difference
    sum_configs p sum_i f(e_i) <psi_i|dH/dlambda|psi_i> (the constant -2 kappa term cancels); the pair table entry by four
    Gauss-Legendre path points (epi_pair_u) equals the same quadrature of the brute-force difference; and that quadrature
    converges to the second difference of the brute-force grand potentials Omega(c, s) - Omega(c0, s) - Omega(c, s0) +
    Omega(c0, s0) (eight points: the thermodynamic-integration identity the driver rests on)."""
    eng = _engine()
    from asr_amd import host
    from exact import exact_pair_domega, exact_pair_omega, pair_hamiltonians
    from test_oracle_exact import _ring_pair_setup
    model, op, par, _, _, partner = _ring_pair_setup()

    def sites_of(lam):
        pt = EP.epi_path(par[1], par[2], lam)
        return {k: np.stack([getattr(p, k) for p in par] + [getattr(pt, k)]) for k in ("c", "d", "o", "e")}

    def exact_diff(lam):
        return (exact_pair_domega(model, sites_of, lam, partner, 2, mu, kt)
                - exact_pair_domega(model, sites_of, lam, partner, 1, mu, kt))

    eig = np.concatenate([np.linalg.eigvalsh(h) for lam in (0.0, 1.0) for s in (1, 2)
                          for h, _ in pair_hamiltonians(model, sites_of(lam), partner, s, 4)])
    mu, kt = float(np.quantile(eig, 0.45)), 0.01
    zs, ws = host.contour_fermi(float(eig.min()) - 0.3, mu, kt, host.fermi_npole(kt), 96, 16)
    vec, chi, lo = EP.epi_adapted(np.eye(9)[None])
    G = eng.GpuOperator(op)
    G.set_depth_tol(0.0, "none")
    nlev, nb = 400, 9 * zs.size  # the ring's chains close well before 400 levels; one call per species and path point
    print(f"epi TI ring: {zs.size} contour energies x 9 start vectors, {nlev} levels", flush=True)
    lam = 0.37
    pt = EP.epi_path(par[1], par[2], lam)
    d_gpu = (EP.epi_domega(G, par, pt, 2, zs, ws, vec, chi, lo, nlev, max_blocks=nb)
             - EP.epi_domega(G, par, pt, 1, zs, ws, vec, chi, lo, nlev, max_blocks=nb))
    d_ex = exact_diff(lam)
    assert abs(d_gpu - d_ex) < 1e-6 * max(1.0, abs(d_ex)), (d_gpu, d_ex)
    u_gpu = EP.epi_pair_u(G, par, 1, 2, 2, 1, zs, ws, vec, chi, lo, nlev, nlam=4, max_blocks=nb)
    x, w = host.gauss_legendre(4)
    u_quad = sum(0.5 * wk * exact_diff(0.5 * (1.0 + xk)) for xk, wk in zip(x, w))
    assert abs(u_gpu - u_quad) < 1e-6, (u_gpu, u_quad)
    om = {(lm, s): exact_pair_omega(model, sites_of(lm), partner, s, 4, mu, kt) for lm in (0.0, 1.0) for s in (1, 2)}
    u_ex = om[1.0, 2] - om[0.0, 2] - om[1.0, 1] + om[0.0, 1]
    x8, w8 = host.gauss_legendre(8)
    u8 = sum(0.5 * wk * exact_diff(0.5 * (1.0 + xk)) for xk, wk in zip(x8, w8))
    assert abs(u8 - u_ex) < 2e-5, (u8, u_ex)
    assert abs(u_quad - u_ex) > 10 * abs(u8 - u_ex)  # the path quadrature converges (4 points are coarse here)
