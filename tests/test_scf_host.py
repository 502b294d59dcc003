"""Host side of the SCF contour pass (asr_amd.scf; contour, ed_moments, the task share of chain_worker,
src/asr_class_scf.f90:797-877, 1093-1157, 1414-1432; contour_nodes / contour_nodes_fermi, src/asr_contour.f90).

Pinned by the mathematics of the contours: for G(z) = sum_i w_i / (z - e_i) (poles on the real axis)
    -(1/pi) Im sum_j w_j (z_j - E_nu)^k G(z_j) = sum_{e_i < E_F} w_i (e_i - E_nu)^k     (closed arc, semicircle),
    -(1/pi) Im sum_j w_j G(z_j) = sum_i w_i f(e_i)                                      (Fermi contour),
and the DOS at the probe E_F + i eta is -(1/pi) Im G there.  Each (start orbital, component, spin) gets its own pole set,
so a transposed index or a wrong orbital weight fails.
"""
import numpy as np
import pytest

from asr_amd import host
from asr_amd import scf


def test_gauss_legendre_matches_numpy():
    for n in (1, 4, 16, 33):
        x, w = host.gauss_legendre(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.max(np.abs(x - xr)) < 1e-14 and np.max(np.abs(w - wr)) < 1e-14


@pytest.mark.parametrize("eta_end", [None, 0.05])
def test_semicircle_counts_enclosed_poles(eta_end):
    """contour_nodes: a pole inside the semicircle counts 1, one outside 0; with eta_end the arc stops at z_end above the
    axis (Im z_end = eta_end for eta_end < r / 2) and the sum is the path integral (arg(E_bot - e) - arg(z_end - e)) / pi."""
    ebot, ef = -1.0, 0.2
    z, w = host.contour_nodes(ebot, ef, 48, eta_end)
    c, r = 0.5 * (ebot + ef), 0.5 * (ef - ebot)
    z_end = complex(ef, 0.0) if eta_end is None else complex(c + np.sqrt(r * r - eta_end ** 2), eta_end)
    for e in (-0.6, -0.3, 0.6, -1.5):
        want = (np.angle(ebot - e + 0j) - np.angle(z_end - e)) / np.pi
        if eta_end is None:
            want = 1.0 if ebot < e < ef else 0.0
        v = -np.imag(np.sum(w / (z - e))) / np.pi
        assert abs(v - want) < 1e-9, (e, v, want)
    if eta_end is None:
        # polynomial integrand: the exact real-axis integral
        v = -np.imag(np.sum(w * z ** 3)) / np.pi
        assert abs(v) < 1e-12


def test_fermi_contour_reproduces_the_fermi_function():
    """contour_nodes_fermi: -(1/pi) Im sum_j w_j / (z_j - e) = f(e) from far below mu across the step (kT = 1 mRy, 16
    poles); node count fermi_nodes_count."""
    kt, mu = 0.001, 0.05
    npl = host.fermi_npole(kt)
    assert npl == 16 and host.fermi_npole(1e-6) == 40 and host.fermi_npole(1.0) == 2
    z, w = host.contour_fermi(-1.0, mu, kt, npl, 64, 16)
    assert z.size == host.fermi_nodes_count(npl, 64, 16) == 4 + 64 + 16 + 8 + 16
    assert np.all(z.imag > 0.0)
    for e in (-0.6, -0.2, mu - 0.01, mu - 0.002, mu, mu + 0.001, mu + 0.005, mu + 0.02):
        f = 1.0 / (np.exp((e - mu) / kt) + 1.0)
        v = -np.imag(np.sum(w / (z - e))) / np.pi
        assert abs(v - f) < 2e-7, (e, v, f)


def test_contour_pass_nodes():
    enu = np.array([[[-0.1, 0.05, 0.0]], [[0.02, -0.2, 0.1]]])
    cp = scf.contour(0.3, enu, ebot_offset=1.0, ncont=20, ed_eta=0.02, fermi_kt=0.0)
    assert cp.nq == 28 and cp.zq.size == 29
    assert cp.zq[-1] == complex(0.3, 0.02) and cp.wq[-1] == 0
    assert abs(cp.zq[0].real - (-1.2)) < 0.05  # the arc starts near E_bot = min E_nu - ebot_offset
    cpf = scf.contour(0.3, enu, fermi_kt=0.002, ncont=24)
    assert cpf.nq == host.fermi_nodes_count(host.fermi_npole(0.002), 24, 16)


def test_task_share_partitions_and_rotates():
    """chain_worker's cyclic rule: every task has exactly one owner; with 2 ranks and an even number of energies the
    owner of an energy alternates between consecutive (orbital, component, spin) groups."""
    ntg, ngroup = 6, 7
    for nr in (1, 2, 3, 5):
        owners = [scf.task_owner(t, ntg, nr) for t in range(ntg * ngroup)]
        assert set(owners) <= set(range(nr))
        counts = np.bincount(owners, minlength=nr)
        assert counts.max() - counts.min() <= ngroup  # balanced to a group
    own = np.array([scf.task_owner(t, ntg, 2) for t in range(ntg * ngroup)]).reshape(ngroup, ntg)
    assert np.all(own[1:] == 1 - own[:-1])


def _pole_model(rng, nz_poles=3):
    e = rng.uniform(-0.8, 0.6, nz_poles)
    w = rng.uniform(0.2, 1.0, nz_poles)
    return e, w / w.sum()


@pytest.mark.parametrize("fermi", [False, True])
def test_ed_moments_edc_on_pole_greens(fermi):
    """ed_moments (route edc) on gz = sum_i w_i / (z - e_i) with a pole set per (orbital, component, spin): the moments
    are the occupied sums with the start-orbital weights, the DOS at E_F is -Im G(E_F + i eta) / pi, no causality
    failures; a G with the wrong sign of Im is counted."""
    rng = np.random.default_rng(3)
    efv, eta = 0.1, 0.01
    classes = [scf.ChainClass(op=None, iorbk=[0, 2, 4, 7], worbk=[1, 3, 3, 2]),
               scf.ChainClass(op=None, iorbk=list(range(9)))]
    comps = [scf.Component(1), scf.Component(0), scf.Component(0)]
    nsp = 2
    enu = rng.uniform(-0.1, 0.1, (3, nsp, 3))
    cp = scf.contour(efv, enu, ebot_offset=1.1, ncont=48, ed_eta=eta, fermi_kt=0.002 if fermi else 0.0)
    nz = cp.zq.size
    gz = np.zeros((nz, 9, 3, nsp), np.complex128)
    want_mo = np.zeros((3, nsp, 3, 3))
    want_dos = np.zeros((3, nsp, 3))
    kt = 0.002
    for gg in range(nsp):
        for pp, c in enumerate(comps):
            cl = classes[c.pcl]
            for o, (wo, ll) in enumerate(zip(cl.worbk, cl.lorbk)):
                e, w = _pole_model(rng)
                e = e[np.abs(e - efv) > 0.05] if not fermi else e
                w = w[:e.size]
                gz[:, o, pp, gg] = np.sum(w[None, :] / (cp.zq[:, None] - e[None, :]), axis=1)
                occ = (1.0 / (np.exp((e - efv) / kt) + 1.0)) if fermi else (e < efv).astype(float)
                for kk in range(3):
                    want_mo[pp, gg, kk, ll] += wo * np.sum(w * occ * (e - enu[pp, gg, ll]) ** kk)
                want_dos[pp, gg, ll] += -wo * np.sum(w / (complex(efv, eta) - e)).imag / np.pi
    cg = scf.ChainGreens(gz, None, 0, 0, 0.0)
    r = scf.ed_moments(cp, cg, classes, comps, enu)
    tol = 2e-6 if fermi else 1e-8
    assert np.max(np.abs(r.mo - want_mo)) < tol, np.max(np.abs(r.mo - want_mo))
    assert np.max(np.abs(r.dosv - want_dos)) < 1e-10
    assert np.array_equal(r.qc, r.mo[:, :, 0, :])
    assert r.npsd == 0 and r.nnode == (9 + 4 + 4) * nsp * int(np.count_nonzero(cp.zq.imag > 0))
    bad = gz.copy()
    bad[3, 1, 2, 0] = np.conj(bad[3, 1, 2, 0])
    assert scf.ed_moments(cp, scf.ChainGreens(bad, None, 0, 0, 0.0), classes, comps, enu).npsd == 1


def test_ed_moments_ed3_psd_and_block_moments():
    """ed_moments (route ed3): blocks from a 2x2 pole model whose -Im block is positive semidefinite by construction
    (q0 = sum |u_i|^2 / (z - e_i), q1 = sum u_i v_i / (z - e_i), q2 = sum |v_i|^2 / (z - e_i) in the combination with o);
    the moments of q2 and q1 + o q2 are the occupied sums; npsd = 0; a block with q1 too large is counted."""
    rng = np.random.default_rng(4)
    efv = 0.05
    classes = [scf.ChainClass(op=None, iorbk=[0, 3])]
    comps = [scf.Component(0)]
    par = [dict(o=rng.uniform(-0.05, 0.05, (1, 3)), p=rng.uniform(0.0, 0.01, (1, 3)))]
    enu = np.zeros((1, 1, 3))
    cp = scf.contour(efv, [-0.2], ebot_offset=1.0, ncont=48, ed_eta=0.01, fermi_kt=0.0)
    nz = cp.zq.size
    gb = np.zeros((nz, 2, 1, 1, 4), np.complex128)
    want2 = np.zeros(3)
    want1 = np.zeros(3)
    for o, ll in enumerate(classes[0].lorbk):
        e = np.array([-0.5, -0.1, 0.4])
        u = rng.uniform(0.2, 1.0, 3)
        v = rng.uniform(-1.0, 1.0, 3)
        ol = par[0]["o"][0, ll]
        r = 1.0 / (cp.zq[:, None] - e[None, :])
        q2 = r @ (v * v)
        q1 = r @ (u * v) - ol * q2        # gb1 with q1 + o q2 = sum u v / (z - e)
        q0 = r @ (u * u) - 2.0 * ol * (r @ (u * v)) + ol ** 2 * q2   # gb0 with gb0 + 2 o gb1 + o^2 gb2 = sum u u / (z - e)
        gb[:, o, 0, 0, :3] = np.stack([q0, q1, q2], axis=1)
        occ = e < efv
        want2[ll] += np.sum(v * v * occ)
        want1[ll] += np.sum(u * v * occ)
    cg = scf.ChainGreens(gb[..., 0], gb, 0, 0, 0.0)
    res = scf.ed_moments(cp, cg, classes, comps, enu, params=par)
    for ll in (0, 1):
        assert abs(res.mo[0, 0, 2, ll] - want2[ll]) < 1e-8
        assert abs(res.mo[0, 0, 1, ll] - want1[ll]) < 1e-8
    assert res.npsd == 0 and res.psd_max <= 1.0 + 1e-12 and res.asym_max == 0.0
    gb2 = gb.copy()
    gb2[5, 0, 0, 0, 1] *= 50.0
    assert scf.ed_moments(cp, scf.ChainGreens(gb2[..., 0], gb2, 0, 0, 0.0), classes, comps, enu, params=par).npsd >= 1
