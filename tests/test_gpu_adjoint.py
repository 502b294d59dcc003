"""GPU parity of the adjoint: the engine's gradient of a loss of the chain coefficients against central finite
differences of the oracle's forward chains (h = 1e-5), checkpoint-interval invariance, and the torch.autograd harness
against finite differences of the whole contour objective.

Bar: gradients to 1e-6 relative (the finite-difference error at h = 1e-5 is O(h^2) ~ 1e-10 of the curvature plus
rounding ~ 1e-11); the forward coefficients to 1e-9 relative as the other chain tests.
"""
import numpy as np
import pytest
import torch

from asr_amd import adjoint as AD
from asr_amd import host
from asr_amd import workloads as W

pytestmark = pytest.mark.gpu


def _engine():
    from asr_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    return engine


@pytest.fixture(scope="module")
def wl():
    return W.ternary_adjoint(L=4, levels=2, nchains=12, depth=12, seed=41)


def _oracle_loss(O, wl, theta, la, lb, nlev):
    mem = AD.members_at(wl.members, theta, wl.info["c_offset"])
    tot = 0.0
    for j, m in enumerate(mem):
        a, b, n = O.lanczos(m, m.l0, nlev, tiny=0.0)
        assert n == nlev
        tot += float(la[j] @ a + lb[j] @ b)
    return tot


def test_adjoint_matches_fd_of_oracle(wl):
    eng = _engine()
    from oracle.oracle import OracleOp
    nlev = int(wl.nlev[0])
    nch = len(wl.members)
    rng = np.random.default_rng(5)
    la = rng.standard_normal((nch, nlev))
    lb = rng.standard_normal((nch, nlev))
    theta = wl.info["theta"]
    G = eng.GpuOperator(wl.op)
    mem = AD.members_at(wl.members, theta, wl.info["c_offset"])
    a, b, g, (fms, ams) = G.lanczos_adjoint(mem, nlev, la, lb, ckpt=5)
    O = OracleOp(wl.op)
    for j in (0, 5, 11):
        ar, br, _ = O.lanczos(mem[j], mem[j].l0, nlev, tiny=0.0)
        assert np.max(np.abs(a[j] - ar)) < 1e-9 * np.max(np.abs(ar))
        assert np.max(np.abs(b[j] - br)) < 1e-9 * np.max(np.abs(br))
    grad = g.sum(axis=0)
    h = 1e-5
    fd = np.zeros_like(theta)
    for idx in np.ndindex(*theta.shape):
        tp = theta.copy(); tp[idx] += h
        tm = theta.copy(); tm[idx] -= h
        fd[idx] = (_oracle_loss(O, wl, tp, la, lb, nlev) - _oracle_loss(O, wl, tm, la, lb, nlev)) / (2 * h)
    scale = np.max(np.abs(fd))
    assert scale > 1e-3
    assert np.max(np.abs(grad - fd)) < 1e-6 * scale, np.max(np.abs(grad - fd)) / scale


def test_adjoint_per_chain_and_checkpoint_interval(wl):
    """Per-chain gradients are independent of the checkpoint interval (the recomputation reproduces the forward's
    vectors with the forward's own kernels), and a loss on one chain gives gradient only to that chain."""
    eng = _engine()
    nlev = int(wl.nlev[0])
    nch = len(wl.members)
    rng = np.random.default_rng(6)
    la = rng.standard_normal((nch, nlev))
    lb = rng.standard_normal((nch, nlev))
    mem = AD.members_at(wl.members, wl.info["theta"], wl.info["c_offset"])
    G = eng.GpuOperator(wl.op)
    _, _, g8, _ = G.lanczos_adjoint(mem, nlev, la, lb, ckpt=8)
    for ck in (1, 3, 12):
        _, _, gk, _ = G.lanczos_adjoint(mem, nlev, la, lb, ckpt=ck)
        assert np.max(np.abs(gk - g8)) < 1e-13 * np.max(np.abs(g8)), ck
    la1 = np.zeros_like(la); la1[4] = la[4]
    lb1 = np.zeros_like(lb); lb1[4] = lb[4]
    _, _, g1, _ = G.lanczos_adjoint(mem, nlev, la1, lb1)
    assert np.max(np.abs(g1[4] - g8[4])) < 1e-12 * np.max(np.abs(g8[4]))
    assert np.all(g1[np.arange(nch) != 4] == 0.0)


def test_torch_harness_contour_objective(wl):
    """L(theta) = band energy + charge penalty of every chain on a closed contour: torch.autograd through the engine's
    adjoint against central finite differences of L evaluated with the engine's forward."""
    eng = _engine()
    nlev = int(wl.nlev[0])
    nch = len(wl.members)
    G = eng.GpuOperator(wl.op)
    z, w = host.contour_closed(-1.6, -0.05, 0.6, 24, 8)
    zt, wt = torch.from_numpy(z), torch.from_numpy(w)
    weight = torch.from_numpy(np.asarray(wl.info["conc"])[[(j // 9) % 3 for j in range(nch)]] / 9.0)
    q0 = torch.full((nch,), 0.5, dtype=torch.float64)
    off = wl.info["c_offset"]

    def L(th):
        a, b = AD.chain_coefficients(th, G, wl.members, nlev, off)
        return AD.contour_objective(a, b, zt, wt, weight, q0, lam=3.0)

    theta = torch.tensor(wl.info["theta"], requires_grad=True)
    loss = L(theta)
    loss.backward()
    grad = theta.grad.numpy()
    h = 1e-5
    fd = np.zeros_like(grad)
    with torch.no_grad():
        for idx in np.ndindex(*grad.shape):
            tp = wl.info["theta"].copy(); tp[idx] += h
            tm = wl.info["theta"].copy(); tm[idx] -= h
            fd[idx] = (float(L(torch.from_numpy(tp))) - float(L(torch.from_numpy(tm)))) / (2 * h)
    scale = np.max(np.abs(fd))
    assert scale > 1e-4
    assert np.max(np.abs(grad - fd)) < 1e-6 * scale, np.max(np.abs(grad - fd)) / scale


def test_torch_harness_dos_gradient(wl):
    """d(DOS)/d(theta): a weighted sum of the chains' densities of states at seven energies (broadening 0.05) through
    torch.autograd and the engine's adjoint against central finite differences of the same sum with the engine's
    forward."""
    eng = _engine()
    nlev = int(wl.nlev[0])
    nch = len(wl.members)
    G = eng.GpuOperator(wl.op)
    e = np.linspace(-1.2, 0.4, 7)
    rng = np.random.default_rng(11)
    v = torch.from_numpy(rng.uniform(0.5, 1.5, (nch, e.size)))
    off = wl.info["c_offset"]

    def L(th):
        a, b = AD.chain_coefficients(th, G, wl.members, nlev, off)
        return (v * AD.chain_dos(a, b, e, 0.05)).sum()

    theta = torch.tensor(wl.info["theta"], requires_grad=True)
    L(theta).backward()
    grad = theta.grad.numpy()
    h = 1e-5
    fd = np.zeros_like(grad)
    with torch.no_grad():
        for idx in np.ndindex(*grad.shape):
            tp = wl.info["theta"].copy(); tp[idx] += h
            tm = wl.info["theta"].copy(); tm[idx] -= h
            fd[idx] = (float(L(torch.from_numpy(tp))) - float(L(torch.from_numpy(tm)))) / (2 * h)
    scale = np.max(np.abs(fd))
    assert scale > 1e-4
    assert np.max(np.abs(grad - fd)) < 1e-6 * scale, np.max(np.abs(grad - fd)) / scale


def test_real_chains_beyond_2gb_vectors():
    """The adjoint workload at full size: 256 real members on 127,517 nodes, vectors of 2.35 GB (beyond 2^31 bytes).  The
    member-tile kernel (tile 16; unsigned 32-bit buffer offsets) gives the chains of the one-node-per-wave kernel (tile
    64) and of the oracle (two chains, one at each end of the batch)."""
    eng = _engine()
    from oracle.oracle import OracleOp
    wa = W.ternary_adjoint()
    mem = AD.members_at(wa.members, wa.info["theta"], wa.info["c_offset"])
    assert (wa.op.nn + 1) * 9 * len(mem) * 8 > 2 ** 31
    nlev = 6
    print(f"real chains beyond 2 GB: {wa.op.nn} nodes, {len(mem)} chains x {nlev}, tiles 16 and 64", flush=True)
    G = eng.GpuOperator(wa.op)
    res = {}
    for tile in (16, 64):
        G.set_tile(tile)
        res[tile] = G.lanczos(mem, nlev, tiny=0.0)
    a16, b16, mu16 = res[16][:3]
    a64, b64, mu64 = res[64][:3]
    assert np.array_equal(mu16, mu64)
    assert np.max(np.abs(a16 - a64)) < 1e-12 and np.max(np.abs(b16 - b64)) < 1e-12
    O = OracleOp(wa.op)
    for j in (0, len(mem) - 1):
        ar, br, n = O.lanczos(mem[j], mem[j].l0, nlev, tiny=0.0)
        assert np.max(np.abs(a16[j, :n] - ar[:n])) < 1e-9 * np.max(np.abs(ar[:n]))
        assert np.max(np.abs(b16[j, :n] - br[:n])) < 1e-9 * np.max(np.abs(br[:n]))


def test_adjoint_fused_and_separate_forms(wl):
    """The reverse sweep with the recomputed level u_k+1 and the combination u_bar_k written by the operator's last pass
    (level form 0 on the member-tile pass, tiles 8 and 16) against the separate passes (form 1) and against the
    one-node-per-wave kernel (tile 64, which keeps the separate passes): the same forward coefficients and gradients to
    rounding."""
    eng = _engine()
    nlev = int(wl.nlev[0])
    nch = len(wl.members)
    rng = np.random.default_rng(8)
    la = rng.standard_normal((nch, nlev))
    lb = rng.standard_normal((nch, nlev))
    mem = AD.members_at(wl.members, wl.info["theta"], wl.info["c_offset"])
    G = eng.GpuOperator(wl.op)
    res = {}
    for form, tile in ((0, 16), (1, 16), (0, 8), (1, 8), (0, 64)):
        G.set_level_form(form)
        G.set_tile(tile)
        res[(form, tile)] = G.lanczos_adjoint(mem, nlev, la, lb, ckpt=4)
    G.set_level_form(0)
    G.set_tile(0)
    a0, b0, g0, _ = res[(0, 16)]
    for key, (a, b, g, _) in res.items():
        assert np.max(np.abs(a - a0)) < 1e-12 * np.max(np.abs(a0)), key
        assert np.max(np.abs(b - b0)) < 1e-12 * np.max(np.abs(b0)), key
        assert np.max(np.abs(g - g0)) < 1e-11 * np.max(np.abs(g0)), key
