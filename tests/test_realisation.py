"""Independent pins of the disordered supercell realisation that the KPM moments run on (h_apply / real_apply,
src/reference/asr_lmto_realization.f90:120-191; kpm_moments, src/reference/asr_kpm.f90:39-65).

* The realisation as a dense matrix built here from its definition, h = C + D S D with S the on-site block plus one
  bond block per neighbour slot (the blocks of a periodic image added where the box wraps), C and D diagonal per site
  species and orbital, real_apply = E + h - h O h: the oracle's and the GPU's applications equal it to 1e-13 and h is
  symmetric (test/test_lmto_real.f90:1, (1)).
* Chebyshev moments mu_n = <r|T_n(H~)|r> / dim with the Rademacher vectors of the engine, from the eigen-decomposition
  of the dense matrix and T_n(cos t) = cos(n t): the GPU's moments (direct and by moment doubling) to 1e-12.
* The ordered limit (test/test_lmto_real.f90:2-3): with every site of one species the Lanczos chain of the realisation
  from a site orbital equals the chain of the ordered augmented-space operator (one node per site, unreduced) from the
  same start, 1e-10.
"""
import numpy as np
import pytest

from asr_amd import augcls as A
from asr_amd import structure as S
from asr_amd import workloads as W
from oracle.oracle import OracleOp, OracleReal, rademacher


def dense_h(nbr, species, par, block, block0):
    n = nbr.shape[0]
    s = np.zeros((9 * n, 9 * n))
    for i in range(n):
        s[9 * i:9 * i + 9, 9 * i:9 * i + 9] += block0
        for k in range(nbr.shape[1]):
            j = nbr[i, k]
            s[9 * i:9 * i + 9, 9 * j:9 * j + 9] += block[k]
    c = par[0][species].reshape(-1)
    d = par[1][species].reshape(-1)
    return np.diag(c) + d[:, None] * s * d[None, :]


def supercell(L, seed, oe=False):
    g, species, par, block, block0 = W.kpm_supercell(L=L, seed=seed)
    if oe:
        rng = np.random.default_rng(seed + 1)
        par = par.copy()
        par[2] = rng.uniform(-0.05, 0.05, par[2].shape)
        par[3] = rng.uniform(-0.1, 0.1, par[3].shape)
    return g, species, par, block, block0


def real_dense(g, species, par, block, block0):
    h = dense_h(g.nbr, species, par, block, block0)
    o = par[2][species].reshape(-1)
    e = par[3][species].reshape(-1)
    return h, np.diag(e) + h - h @ (o[:, None] * h)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def exact_moments(h, c, s, nmom, nvec, seed):
    """mu_per[v, n] = <r_v|T_n((h - c) / s)|r_v> / dim from the eigen-decomposition of h."""
    lam, vec = np.linalg.eigh(h)
    x = (lam - c) / s
    assert np.max(np.abs(x)) < 1.0
    th = np.arccos(x)
    dim = h.shape[0]
    out = np.zeros((nvec, nmom))
    for v in range(nvec):
        r = np.array([rademacher(seed, v, q) for q in range(dim)])
        w = (vec.T @ r) ** 2
        out[v] = np.cos(np.outer(np.arange(nmom), th)) @ w / dim
    return out


@pytest.mark.parametrize("L", [2, 3])
def test_oracle_realisation_equals_dense(L):
    g, species, par, block, block0 = supercell(L, 40 + L, oe=True)
    h, y_mat = real_dense(g, species, par, block, block0)
    assert np.max(np.abs(h - h.T)) < 1e-14 * np.max(np.abs(h))
    assert len(np.unique(species)) == 3
    O = OracleReal(g.nbr, species, par[0], par[1], par[2], par[3], block, block0)
    rng = np.random.default_rng(L)
    x = rng.standard_normal((3, 9 * g.n))
    for j in range(3):
        assert rel(O.apply(x[j]), y_mat @ x[j]) < 1e-13
    # a transposed bond block or d applied on one side only is not this matrix
    bad = dense_h(g.nbr, species, par, np.transpose(block, (0, 2, 1)), block0)
    assert rel(bad, h) > 1e-3
    cdiag = np.diag(par[0][species].reshape(-1))
    d = par[1][species].reshape(-1)
    one_side = cdiag + (h - cdiag) / d[None, :]  # C + D S
    assert rel(one_side, h) > 1e-3


def test_oracle_moments_equal_exact():
    g, species, par, block, block0 = supercell(2, 7)
    h = dense_h(g.nbr, species, par, block, block0)
    lam = np.linalg.eigvalsh(h)
    c, s = 0.5 * (lam[0] + lam[-1]), 0.55 * (lam[-1] - lam[0])
    O = OracleReal(g.nbr, species, par[0], par[1], par[2], par[3], block, block0)
    mu, per = O.kpm_moments(c, s, 40, 3, 9, per_vector=True)
    ex = exact_moments(h, c, s, 40, 3, 9)
    assert np.max(np.abs(per - ex)) < 1e-12
    assert np.max(np.abs(mu - ex.mean(axis=0))) < 1e-12


def _ordered_pair(L=4, levels=5, seed=3):
    """All-A fcc supercell realisation and the ordered augmented-space operator of the same graph and blocks."""
    g, block, block0 = W.graph_and_blocks("fcc", L)
    rng = np.random.default_rng(seed)
    c, d = W.draw_cd(rng, 1)
    species = np.zeros(g.n, np.uint8)
    par = np.zeros((4, 1, 9))
    par[0], par[1] = c, d
    cls = np.zeros(g.n, np.int32)
    perm = np.arange(g.n, dtype=np.int32)[None, :]
    op = A.build_operator(g, cls, [1], [[0]], [[1.0]], perm, np.eye(9)[None], 0, levels, maxrank=1,
                          block=block[None], block0=block0[None], nsp=1)
    z = np.zeros((1, 9))
    members = [A.Member(l0=l0, c=c, d=d, o=z, e=z) for l0 in range(9)]
    return g, species, par, block, block0, op, members


def lanczos_np(apply, dim, start, nlev):
    """Plain Lanczos with full reorthogonalisation on a matrix-free operator (numpy)."""
    q = np.zeros(dim)
    q[start] = 1.0
    qs = [q]
    a = np.zeros(nlev)
    b = np.zeros(nlev)
    qp = np.zeros(dim)
    for k in range(nlev):
        w = apply(q)
        a[k] = q @ w
        w = w - a[k] * q - (b[k - 1] * qp if k else 0.0)
        for _ in range(2):
            w = w - np.stack(qs).T @ (np.stack(qs) @ w)
        b[k] = np.linalg.norm(w)
        qp, q = q, w / b[k]
        qs.append(q)
    return a, b


def test_oracle_ordered_limit_equals_augmented_space():
    g, species, par, block, block0, op, members = _ordered_pair()
    Or = OracleReal(g.nbr, species, par[0], par[1], par[2], par[3], block, block0)
    Oa = OracleOp(op)
    for m in members[::2]:
        a1, b1 = lanczos_np(Or.apply, 9 * g.n, m.l0, 5)
        a2, b2, n = Oa.lanczos(m, m.l0, 5, reorth=True)
        assert n == 5
        assert max(np.max(np.abs(a1 - a2)), np.max(np.abs(b1[:4] - b2[:4]))) < 1e-10


# ------------------------------------------------------------------------------------------------ GPU
def _engine():
    from asr_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    return engine


@pytest.mark.gpu
@pytest.mark.parametrize("L,tile", [(2, 1), (2, 2), (3, 1), (3, 2), (3, 3)])
def test_gpu_realisation_equals_dense(L, tile):
    eng = _engine()
    g, species, par, block, block0 = supercell(L, 40 + L)
    h = dense_h(g.nbr, species, par, block, block0)
    R = eng.GpuRealization(g.nbr, species, par, block, block0)
    R.set_tile(tile)
    x = np.random.default_rng(L).standard_normal((5, 9 * g.n))
    y = R.apply(x)
    for j in range(5):
        assert rel(y[j], h @ x[j]) < 1e-13, j


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [2, 4])
@pytest.mark.parametrize("doubling", [False, True])
def test_gpu_moments_equal_exact(doubling, tile):
    eng = _engine()
    g, species, par, block, block0 = supercell(2, 7)
    h = dense_h(g.nbr, species, par, block, block0)
    lam = np.linalg.eigvalsh(h)
    c, s = 0.5 * (lam[0] + lam[-1]), 0.55 * (lam[-1] - lam[0])
    R = eng.GpuRealization(g.nbr, species, par, block, block0)
    R.set_tile(tile)
    R.set_doubling(doubling)
    ex = exact_moments(h, c, s, 40, 5, 9)
    res = R.kpm_moments(c, s, 40, 5, 9, per_vector=True)
    assert np.max(np.abs(res["mu_per"] - ex)) < 1e-12
    assert np.max(np.abs(res["mu"] - ex.mean(axis=0))) < 1e-12


@pytest.mark.gpu
def test_gpu_ordered_limit_equals_augmented_space():
    """The GPU realisation (Lanczos in numpy over asr_gpu_real_apply) and the GPU augmented-space chains
    (asr_gpu_lanczos_real on the ordered operator) give the same 5 levels, 1e-10."""
    eng = _engine()
    g, species, par, block, block0, op, members = _ordered_pair()
    R = eng.GpuRealization(g.nbr, species, par, block, block0)
    G = eng.GpuOperator(op)
    a2, b2, mu, _ = G.lanczos(members, 5)
    for j, m in enumerate(members):
        a1, b1 = lanczos_np(lambda v: R.apply(v)[0], 9 * g.n, m.l0, 5)
        assert mu[j] == 5
        assert max(np.max(np.abs(a1 - a2[j, :5])), np.max(np.abs(b1[:4] - b2[j, :4]))) < 1e-10, j
