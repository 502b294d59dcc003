"""GPU against the closed forms and golden recipes of the reference (through the C ABI): KPM on the periodic sc band
against its exact k-space spectrum (test/test_kpm.f90), the ordered nearest-neighbour band (a_0 = 0, b_0^2 = z), and the
DOS of the plain binary workload (9 real chains x 20 levels, Beer-Pettifor terminator, 2001 energies) against the
oracle's chains: integrated DOS to 1e-6 states per atom."""
import numpy as np
import pytest

from asr_amd import augcls as A
from asr_amd import host
from asr_amd import structure as S
from asr_amd import workloads as W

from test_golden_cpu import _sc_band, kpm_dos_error

pytestmark = pytest.mark.gpu


def _engine():
    from asr_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    return engine


def test_gpu_kpm_reproduces_exact_sc_spectrum():
    eng = _engine()
    g, spec, par, block, block0 = _sc_band(12)
    R = eng.GpuRealization(g.nbr, spec, par, block, block0)
    for vt, dbl in ((1, False), (2, False), (2, True), (4, False), (4, True)):
        R.set_tile(vt)
        R.set_doubling(dbl)  # moment doubling: two moments per Chebyshev step
        res = R.kpm_moments(0.0, 6.3, 400, 60, seed=5)
        assert abs(res["mu"][0] - 1.0) < 1e-14
        assert kpm_dos_error(res["mu"]) < 2.5e-2, (vt, dbl)
    R.close()


@pytest.mark.parametrize("kind,z", [("sc", 6), ("bcc", 8), ("fcc", 12)])
def test_gpu_ordered_nearest_neighbour_band(kind, z):
    eng = _engine()
    L = 4
    g = S.periodic_graph(kind, L)
    cls = np.zeros(g.n, np.int32)
    perm, mats = S.class_ops(g.coord, 2 * L, cls, "identity")
    nslot = g.nbr.shape[1]
    block = -np.repeat(np.eye(9)[None], nslot, axis=0)
    op = A.build_operator(g, cls, [1], [[0]], [[1.0]], perm, S.orbital_rep(mats), 0, 3, maxrank=1,
                          block=block[None], block0=np.zeros((1, 9, 9)), nsp=1)
    members = [A.Member(l0=l0, c=np.zeros((1, 9)), d=np.ones((1, 9)), o=np.zeros((1, 9)), e=np.zeros((1, 9)))
               for l0 in range(9)]
    G = eng.GpuOperator(op)
    a, b, mu, _ = G.lanczos(members, 3)
    assert np.max(np.abs(a[:, 0])) < 1e-14
    assert np.max(np.abs(b[:, 0] ** 2 - z)) < 1e-12


def test_gpu_dos_of_plain_binary_matches_oracle():
    eng = _engine()
    from oracle.oracle import OracleOp
    wl = W.plain_binary(L=6, levels=3, depth=20, seed=7)
    G = eng.GpuOperator(wl.op)
    a, b, mu, _ = G.lanczos(wl.members, wl.nlev)
    O = OracleOp(wl.op)
    e = np.linspace(-2.5, 2.5, 2001)
    z = e + 0.01j
    for j, m in enumerate(wl.members):
        ar, br, n = O.lanczos(m, m.l0, 20)
        assert mu[j] == n == 20
        rho = []
        for aa, bb in ((a[j], b[j]), (ar, br)):
            ainf, binf, _ = host.bp_asymptote(aa, bb, 20)
            gz = host.terminated_green(aa, bb, 20, ainf, binf, z)
            rho.append(-gz.imag / np.pi)
        nint = [np.concatenate([[0.0], np.cumsum(0.5 * (r[1:] + r[:-1]) * np.diff(e))]) for r in rho]
        assert np.max(np.abs(nint[0] - nint[1])) < 1e-6, j
        assert 0.5 < nint[1][-1] <= 1.0 + 1e-9


def test_gpu_spectral_bounds():
    """spectral_bounds (src/reference/asr_kpm.f90:14-36) on the GPU: on the 12^3 sc band the Ritz extremes of 80 Lanczos
    steps are the exact band edges -6 and 6 (interlacing keeps them inside, convergence to 1e-10); on a random ternary fcc
    supercell the coefficients of each start vector are the oracle's (vector v = the oracle's seed + v) and the bounds are
    the extremes of the oracle's tridiagonal matrices."""
    eng = _engine()
    from oracle.oracle import OracleReal
    g, spec, par, block, block0 = _sc_band(12)
    R = eng.GpuRealization(g.nbr, spec, par, block, block0)
    lo, hi, a, b = R.spectral_bounds(80, margin=0.05, nvec=2, seed=11)
    assert abs(lo + 0.05 + 6.0) < 1e-10 and abs(hi - 0.05 - 6.0) < 1e-10
    R.close()
    g, species, par, block, block0 = W.kpm_supercell(L=8, seed=6)
    R = eng.GpuRealization(g.nbr, species, par, block, block0)
    lo, hi, a, b = R.spectral_bounds(80, margin=0.0, nvec=3, seed=11)
    O = OracleReal(g.nbr, species, par[0], par[1], par[2], par[3], block, block0)
    los, his = [], []
    for v in range(3):
        ao, bo = O.lanczos(80, 11 + v)
        assert np.max(np.abs(a[v, :30] - ao[:30])) < 1e-9 * np.max(np.abs(ao[:30]))
        assert np.max(np.abs(b[v, :30] - bo[:30])) < 1e-9 * np.max(np.abs(bo[:30]))
        ev = host.tridiag_eigvals(ao, bo[:79])
        los.append(ev[0])
        his.append(ev[-1])
    assert abs(lo - min(los)) < 1e-10 and abs(hi - max(his)) < 1e-10
    R.close()
