"""GPU parity: the HIP engine (through the C ABI) against the oracle on the same seeded inputs.

Bar: FP64 chain coefficients a_n, b_n to 1e-9 relative over the levels run, operator applications to 1e-12 relative,
KPM moments to 1e-10 absolute (moments are O(1)).
"""
import numpy as np
import pytest

from asr_amd import augcls as A
from asr_amd import structure as S
from asr_amd import workloads as W

pytestmark = pytest.mark.gpu


def _engine():
    from asr_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    return engine


def _oracle(op):
    from oracle.oracle import OracleOp
    return OracleOp(op)


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300))


@pytest.fixture(scope="module")
def binary_small():
    return W.plain_binary(L=4, levels=2, depth=12, seed=11)


@pytest.fixture(scope="module")
def ternary_small():
    return W.ternary_edc(L=4, levels=2, depth=14, seed=12)


def test_apply_real_members(binary_small):
    eng = _engine()
    wl = binary_small
    op = wl.op
    rng = np.random.default_rng(0)
    # three members: non-uniform d, a frozen-free chain, o != 0 and e != 0 on one of them
    m0 = wl.members[0]
    m1 = wl.members[4]
    m2 = A.Member(l0=6, c=m0.c, d=m0.d, o=rng.uniform(-0.05, 0.05, (2, 9)), e=rng.uniform(-0.1, 0.1, (2, 9)))
    members = [m0, m1, m2]
    x = rng.standard_normal((3, op.dim))
    G = eng.GpuOperator(op)
    y = G.apply(members, x)
    O = _oracle(op)
    for j, m in enumerate(members):
        assert rel(y[j], O.apply(m, x[j])) < 1e-12, f"member {j}"


def test_apply_complex_edc_operator(ternary_small):
    eng = _engine()
    wl = ternary_small
    op = wl.op
    rng = np.random.default_rng(1)
    members = [wl.members[k] for k in (0, 13, 40)]
    x = rng.standard_normal((3, op.dim)) + 1j * rng.standard_normal((3, op.dim))
    G = eng.GpuOperator(op)
    y = G.apply_complex(members, x)
    O = _oracle(op)
    for j, m in enumerate(members):
        ep = m.z * m.pa + m.pb
        ref = O.apply(m, x[j].real) + 1j * O.apply(m, x[j].imag) + O.rand_c(m, ep, x[j])
        assert rel(y[j], ref) < 1e-12, f"member {j}"


def test_lanczos_real_batched_and_plain(binary_small):
    eng = _engine()
    wl = binary_small
    G = eng.GpuOperator(wl.op)
    a, b, mu, _ = G.lanczos(wl.members, wl.nlev, mode=0)
    a1, b1, mu1, _ = G.lanczos(wl.members[:3], wl.nlev[:3], mode=1)
    O = _oracle(wl.op)
    for j, m in enumerate(wl.members):
        ar, br, n = O.lanczos(m, m.l0, int(wl.nlev[j]))
        assert mu[j] == n
        assert rel(a[j, :n], ar[:n]) < 1e-9
        assert rel(b[j, :n], br[:n]) < 1e-9
    # the plain form reduces over a different tile shape (one member per lane): agreement to rounding
    assert rel(a1, a[:3]) < 1e-13 and rel(b1, b[:3]) < 1e-13
    a2, b2, mu2, _ = G.lanczos(wl.members, wl.nlev, mode=0)
    assert np.array_equal(a, a2) and np.array_equal(b, b2), "a batched run must be bitwise reproducible"


def test_lanczos_uneven_depths(binary_small):
    eng = _engine()
    wl = binary_small
    G = eng.GpuOperator(wl.op)
    nlev = np.array([3, 12, 7, 1, 9, 12, 5, 2, 11], np.int32)
    a, b, mu, _ = G.lanczos(wl.members, nlev)
    O = _oracle(wl.op)
    for j, m in enumerate(wl.members):
        ar, br, n = O.lanczos(m, m.l0, int(nlev[j]))
        assert mu[j] == n == nlev[j]
        assert rel(a[j, :n], ar[:n]) < 1e-9
        assert rel(b[j, :n], br[:n]) < 1e-9
        assert np.all(a[j, n:] == 0.0)


@pytest.mark.parametrize("project", [True, False])
def test_edc_chains_batched(ternary_small, project):
    eng = _engine()
    wl = ternary_small
    G = eng.GpuOperator(wl.op)
    idx = [0, 4, 9, 17, 22, 30, 44, 53]
    members = [wl.members[k] for k in idx]
    nlev = np.array([14, 14, 10, 14, 6, 14, 13, 14], np.int32)
    a, b, mu, _ = G.chains_edc(members, nlev, project=project)
    O = _oracle(wl.op)
    for j, m in enumerate(members):
        ep = m.z * m.pa + m.pb
        ar, br, n = O.edc_chain(m, ep, m.l0, int(nlev[j]), project=project)
        assert mu[j] == n
        assert rel(a[j, :n], ar[:n]) < 1e-9, f"chain {j}"
        assert rel(b[j, :n - 1], br[:n - 1]) < 1e-9, f"chain {j}"


def test_edc_plain_equals_batched(ternary_small):
    eng = _engine()
    wl = ternary_small
    G = eng.GpuOperator(wl.op)
    members = wl.members[:4]
    a, b, mu, _ = G.chains_edc(members, 10, mode=0)
    a1, b1, mu1, _ = G.chains_edc(members, 10, mode=1)
    assert np.array_equal(mu, mu1)
    assert rel(a1, a) < 1e-13 and rel(b1, b) < 1e-13


@pytest.mark.parametrize("tile", [1, 2, 3, 4])
def test_kpm_apply_and_moments(tile):
    """Chebyshev steps with one or two vectors per lane (64-vector rows), on 16-vector slices or with the block contraction
    on the FP64 matrix cores (tile 4), against the oracle."""
    eng = _engine()
    from oracle.oracle import OracleReal
    g, species, par, block, block0 = W.kpm_supercell(L=3, seed=21)
    R = eng.GpuRealization(g.nbr, species, par, block, block0)
    R.set_tile(tile)
    O = OracleReal(g.nbr, species, par[0], par[1], par[2], par[3], block, block0)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((3, R.dim))
    y = R.apply(x)
    for j in range(3):
        assert rel(y[j], O.apply(x[j])) < 1e-12
    c, s = 0.05, 2.5
    res = R.kpm_moments(c, s, 48, 5, seed=7, per_vector=True)
    mu_o, per_o = O.kpm_moments(c, s, 48, 5, seed=7, per_vector=True)
    assert np.max(np.abs(res["mu_per"] - per_o)) < 1e-10
    assert np.max(np.abs(res["mu"] - mu_o)) < 1e-10
    assert abs(res["mu"][0] - 1.0) < 1e-14
    # moment doubling (mu_2m = 2 <t_m|t_m> - mu_0, mu_2m-1 = 2 <t_m|t_m-1> - mu_1): the direct moments up to rounding, for
    # an even and an odd count
    R.set_doubling(True)
    for nm in (48, 47, 3):
        rd = R.kpm_moments(c, s, nm, 5, seed=7, per_vector=True)
        assert np.max(np.abs(rd["mu_per"] - per_o[:, :nm])) < 1e-10, nm
    R.set_doubling(False)


@pytest.mark.parametrize("tile", [8, 16, 64])
def test_edc_two_site_types(tile):
    """Two site types (type 1: the bond blocks x 1.1 and the on-site block shifted by 0.01): the operator pass takes each
    node's blocks by its type (the member-tile kernel's LDS rows, not its one-type scalar path) and gives the oracle's
    chains; the type-1 blocks change the chains by about 0.1, so a pass that ignores the type fails."""
    eng = _engine()
    g, block, block0 = W.graph_and_blocks("fcc", 4)
    cls = np.zeros(g.n, np.int32)
    perm, mats = S.class_ops(g.coord, 8, cls, "d2h")
    dmat = S.orbital_rep(mats)
    fl = np.ones(g.n, np.uint8)
    fl[0] = 0
    stype = (np.arange(g.n) % 2).astype(np.int32)
    op = A.reorder_nodes(A.build_operator(g, cls, [3], [[0, 1, 2]], [[0.66, 0.14, 0.2]], perm, dmat, 0, 2, maxrank=1,
                                          flippable=fl, stype=stype, block=np.stack([block, 1.1 * block]),
                                          block0=np.stack([block0, block0 + 0.01 * np.eye(9)]), nsp=3))
    assert op.ntype == 2
    spins = W.edc_params(np.random.default_rng(5), 3, 2)
    members = [W.edc_member(spins[sp], cen, l0, complex(-0.2, 0.15)) for sp in range(2) for cen in range(3)
               for l0 in range(0, 9, 2)]
    nlev = np.array([8 + (k % 4) for k in range(len(members))], np.int32)
    print(f"two site types: {op.nn} nodes, tile {tile}, {len(members)} chains", flush=True)
    G = eng.GpuOperator(op)
    G.set_tile(tile)
    a, b, mu, _ = G.chains_edc(members, nlev)
    O = _oracle(op)
    for j in (0, 7, 18, len(members) - 1):
        m = members[j]
        ar, br, n = O.edc_chain(m, m.z * m.pa + m.pb, m.l0, int(nlev[j]))
        assert mu[j] == n
        assert rel(a[j, :n], ar[:n]) < 1e-9, f"chain {j}"
        assert rel(b[j, :n - 1], br[:n - 1]) < 1e-9, f"chain {j}"


@pytest.mark.parametrize("tile", [8, 16, 64, 65])
def test_edc_member_tiles(ternary_small, tile):
    """Every member tile (several nodes per wave, or one node per wave with 64 members) gives the oracle's chains."""
    eng = _engine()
    wl = ternary_small
    G = eng.GpuOperator(wl.op)
    G.set_tile(tile)
    members = wl.members  # 54 chains: one ragged 64-member tile
    nlev = np.array([8 + (k % 5) for k in range(len(members))], np.int32)
    a, b, mu, _ = G.chains_edc(members, nlev)
    O = _oracle(wl.op)
    for j in (0, 7, 26, 53):
        m = members[j]
        ar, br, n = O.edc_chain(m, m.z * m.pa + m.pb, m.l0, int(nlev[j]))
        assert mu[j] == n
        assert rel(a[j, :n], ar[:n]) < 1e-9, f"chain {j}"
        assert rel(b[j, :n - 1], br[:n - 1]) < 1e-9, f"chain {j}"


def test_real_member_tile_64(binary_small):
    eng = _engine()
    wl = binary_small
    G = eng.GpuOperator(wl.op)
    G.set_tile(64)
    members = wl.members * 5  # 45 real chains
    a, b, mu, _ = G.lanczos(members, 10)
    O = _oracle(wl.op)
    for j in (0, 13, 44):
        ar, br, n = O.lanczos(members[j], members[j].l0, 10)
        assert rel(a[j, :n], ar[:n]) < 1e-9 and rel(b[j, :n], br[:n]) < 1e-9


@pytest.fixture(scope="module")
def l21_small():
    return W.l21_batch(L=4, levels=(2, 2, 2, 1), nchains=40, depth=(5, 12), seed=31)


@pytest.mark.parametrize("mode", [0, 2])
def test_multi_group_chains_with_sro(l21_small, mode):
    """Four sublattice operators (SRO shell classes, per-chain conditional concentrations, uneven depths) in one batched run:
    host-issued streams (mode 0) and one captured HIP graph (mode 2) against the oracle chain by chain."""
    eng = _engine()
    groups = [(eng.GpuOperator(w.op), w.members, w.nlev) for w in l21_small]
    res, ms = eng.chains_multi(groups, mode=mode)
    for (G, members, nlev), (a, b, mu), w in zip(groups, res, l21_small):
        O = _oracle(w.op)
        for j in range(0, len(members), 3):
            m = members[j]
            ar, br, n = O.edc_chain(m, m.z * m.pa + m.pb, m.l0, int(nlev[j]))
            assert mu[j] == n
            assert rel(a[j, :n], ar[:n]) < 1e-9, f"{w.name} chain {j}"
            if n > 1:
                assert rel(b[j, :n - 1], br[:n - 1]) < 1e-9, f"{w.name} chain {j}"


def test_multi_group_graph_replays_agree(l21_small):
    """Two graph-captured runs and the host-issued run give the same coefficients (graph replays hold no stale state)."""
    eng = _engine()
    groups = [(eng.GpuOperator(w.op), w.members, w.nlev) for w in l21_small]
    r0, _ = eng.chains_multi(groups, mode=0)
    r1, _ = eng.chains_multi(groups, mode=2)
    r2, _ = eng.chains_multi(groups, mode=2)
    for (a0, b0, m0), (a1, b1, m1), (a2, b2, m2) in zip(r0, r1, r2):
        assert np.array_equal(m0, m1) and np.array_equal(m1, m2)
        assert np.array_equal(a1, a2) and np.array_equal(b1, b2)
        assert np.array_equal(a0, a1) and np.array_equal(b0, b1)


@pytest.mark.parametrize("mode", [0, 1])
def test_edc_adaptive_depth(mode):
    """Adaptive depth of edc_green (src/asr_edc.f90:121-127): at levels 8, 12, ... the chain terminated with the hull tail;
    a chain leaves at level k >= 12 when |G_k - G_{k-4}| <= depth_tol |G_k|.  The stopping levels are those of the rule
    applied to the chains' own coefficients (host cf_species), the coefficients up to there are the oracle's, and the
    tolerance (the median change at level 12) makes about half the chains leave early."""
    eng = _engine()
    from asr_amd import host
    wl = W.ternary_edc(L=4, levels=2, depth=24, seed=13)
    G = eng.GpuOperator(wl.op)
    members = wl.members[:20]
    nlev = 24
    a0, b0, mu0, _ = G.chains_edc(members, nlev, mode=mode)
    assert np.all(mu0 == nlev)

    def g_at(a, b, k1, z):
        return host.cf_species(a, b, np.zeros(k1), k1, z, "hull")[0]

    ch = np.array([abs(g_at(a0[j], b0[j], 12, m.z) - g_at(a0[j], b0[j], 8, m.z)) / abs(g_at(a0[j], b0[j], 12, m.z))
                   for j, m in enumerate(members)])
    # a tolerance between two distinct changes near the median, well away from both (equivalent orbitals give equal
    # chains, so a tolerance equal to one of the changes would leave the decision to rounding)
    u = np.unique(np.round(ch, 12))
    mid = len(u) // 2
    lo, hi = u[mid - 1], u[mid]
    assert hi > lo * (1 + 1e-6)
    tol = float(np.sqrt(lo * hi))
    G.set_depth_tol(tol, "hull")
    a, b, mu, _ = G.chains_edc(members, nlev, mode=mode)
    G.set_depth_tol(0.0)
    O = _oracle(wl.op)
    for j, m in enumerate(members):
        stop, gprev = nlev, None
        for k1 in range(8, nlev, 4):
            g = g_at(a[j], b[j], k1, m.z)
            if k1 > 8 and abs(g - gprev) <= tol * abs(g):
                stop = k1
                break
            gprev = g
        assert mu[j] == stop, (j, mu[j], stop)
        ar, br, n = O.edc_chain(m, m.z * m.pa + m.pb, m.l0, nlev)
        k = int(mu[j])
        assert rel(a[j, :k], ar[:k]) < 1e-9
        assert rel(b[j, :k - 1], br[:k - 1]) < 1e-9
        assert np.all(a[j, k:] == 0.0)
    assert np.any(mu == 12) and np.any(mu > 12)


@pytest.mark.parametrize("tail,mode", [("hull", 2), ("avg2", 2), ("none", 0), ("bp", 0)])
def test_edc_adaptive_depth_tails_and_graph(tail, mode):
    """The adaptive depth with every tail of cf_species: hull, avg2 and none are tested on the device (the chain
    matrix's eigenvalues by the complex-symmetric QL iteration), so mode 2 stays one HIP graph; bp is tested on the host.
    The stopping levels are those of the rule applied with the host's cf_species (its own eigen-solver) to the chains'
    coefficients."""
    eng = _engine()
    from asr_amd import host
    wl = W.ternary_edc(L=4, levels=2, depth=24, seed=13)
    G = eng.GpuOperator(wl.op)
    members = wl.members[20:44]
    nlev = 24

    def g_at(a, b, k1, z):
        return host.cf_species(a, b, np.zeros(k1), k1, z, tail)[0]

    a0, b0, _, _ = G.chains_edc(members, nlev)
    ch = np.array([abs(g_at(a0[j], b0[j], 12, m.z) - g_at(a0[j], b0[j], 8, m.z)) / abs(g_at(a0[j], b0[j], 12, m.z))
                   for j, m in enumerate(members)])
    u = np.unique(np.round(ch, 12))
    mid = len(u) // 2
    lo, hi = u[mid - 1], u[mid]
    assert hi > lo * (1 + 1e-6)
    tol = float(np.sqrt(lo * hi))
    G.set_depth_tol(tol, tail)
    res, _ = eng.chains_multi([(G, members, np.full(len(members), nlev, np.int32))], mode=mode)
    a, b, mu = res[0]
    G.set_depth_tol(0.0)
    for j, m in enumerate(members):
        stop, gprev = nlev, None
        for k1 in range(8, nlev, 4):
            g = g_at(a[j], b[j], k1, m.z)
            if k1 > 8 and abs(g - gprev) <= tol * abs(g):
                stop = k1
                break
            gprev = g
        assert mu[j] == stop, (j, mu[j], stop)
        k = int(mu[j])
        assert rel(a[j, :k], a0[j, :k]) < 1e-12 and np.all(a[j, k:] == 0.0)
    assert np.any(mu == 12) and np.any(mu > 12)


def test_lanczos_full_reorthogonalisation():
    """lanczos with reorth = .true. on the GPU: the chains of the plain binary workload to depth 40 equal the oracle's
    reorthogonalised chains (1e-9); on the ring the reorthogonalised chain closes after dim steps with the exact spectrum
    of the dense operator (multiplicities included)."""
    eng = _engine()
    from asr_amd import host
    from exact import build_ring_operator, dense_operator, ring_model
    from test_oracle_exact import rotated
    wl = W.plain_binary(L=4, levels=2, depth=40, seed=11)
    G = eng.GpuOperator(wl.op)
    G.set_reorth(True)
    a, b, mu, _ = G.lanczos(wl.members, 40)
    O = _oracle(wl.op)
    for j, m in enumerate(wl.members):
        ao, bo, n = O.lanczos(m, m.l0, 40, reorth=True)
        assert mu[j] == n
        assert rel(a[j, :n], ao[:n]) < 1e-9 and rel(b[j, :n], bo[:n]) < 1e-9, j
    G.close()
    model, par = ring_model()
    model = rotated(model, 2)
    op = build_ring_operator(model, frozen_start=True)
    m = A.Member(l0=4, c=par["c"], d=par["d"], o=par["o"], e=par["e"], frozen=0)
    O = _oracle(op)
    h = dense_operator(lambda x: O.apply(m, x), op.dim)
    ev = np.linalg.eigvalsh(0.5 * (h + h.T))
    G = eng.GpuOperator(op)
    G.set_reorth(True)
    a, b, mu, _ = G.lanczos([m], op.dim + 8, tiny=1.0e-9)
    n = int(mu[0])
    assert n == op.dim and b[0, n - 1] < 1.0e-9
    ritz = host.tridiag_eigvals(a[0, :n], b[0, :n - 1])
    assert np.max(np.abs(np.sort(ritz) - ev)) < 1.0e-9
    G.close()
