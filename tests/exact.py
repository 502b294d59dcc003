"""Independent truths for the oracle: exact configuration averages on tiny rings, dense matrices of the unreduced operator.

The ring cases follow the brute-force checks of the reference (test/test_augn.f90:41-126, test_slow/test_augcls.f90:82-193):
every configuration of the random sites is enumerated, the resolvent of its dense Hamiltonian is solved, and the
configuration average is compared with the chain of the augmented-space operator (exhausted Krylov space).
"""
from __future__ import annotations

import itertools

import numpy as np

from asr_amd import augcls as A
from asr_amd import structure as S


def ring_graph(n: int) -> S.Graph:
    nbr = np.array([[(i + 1) % n, (i - 1) % n] for i in range(n)], np.int32)
    return S.Graph(n=n, nbr=nbr, coord=np.zeros((n, 3), np.int64))


def ring_model(seed: int = 20261111):
    """Ring of 4 sites, three classes: 0 ordered {0}, 1 binary {1, 2} (0.4, 0.6), 2 ternary {0, 2, 3} (0.3, 0.5, 0.2),
    site classes [0, 1, 2, 1]; random symmetric bond blocks (block of slot 1 = transpose of slot 0)."""
    rng = np.random.default_rng(seed)
    block = np.zeros((1, 2, 9, 9))
    block[0, 0] = rng.uniform(-0.25, 0.25, (9, 9))
    block[0, 1] = block[0, 0].T
    b0 = rng.uniform(-0.3, 0.3, (9, 9))
    b0 = np.triu(b0) + np.triu(b0, 1).T
    model = dict(cls=np.array([0, 1, 2, 1], np.int32), nsc=[1, 2, 3], spc=[[0], [1, 2], [0, 2, 3]],
                 conc=[[1.0], [0.4, 0.6], [0.3, 0.5, 0.2]], block=block, block0=b0[None], nsp=4)
    par = dict(c=rng.uniform(-0.4, 0.4, (4, 9)), d=rng.uniform(0.3, 0.7, (4, 9)), o=rng.uniform(-0.2, 0.0, (4, 9)),
               e=rng.uniform(-0.2, 0.2, (4, 9)), pa=rng.uniform(-0.15, 0.15, (4, 9)), pb=rng.uniform(-0.1, 0.1, (4, 9)))
    return model, par


def build_ring_operator(model, frozen_start: bool, maxlevel: int = 80):
    n = model["cls"].size
    g = ring_graph(n)
    fl = np.ones(n, np.uint8)
    if frozen_start:
        fl[0] = 0
    perm = np.arange(n, dtype=np.int32)[None, :]
    return A.build_operator(g, model["cls"], model["nsc"], model["spc"], model["conc"], perm, np.eye(9)[None], 0, maxlevel,
                            flippable=fl, block=model["block"], block0=model["block0"], nsp=model["nsp"])


def dense_s(model) -> np.ndarray:
    n = model["cls"].size
    g = ring_graph(n)
    s = np.zeros((9 * n, 9 * n))
    for i in range(n):
        s[9 * i:9 * i + 9, 9 * i:9 * i + 9] += model["block0"][0]
        for k in range(2):
            j = g.nbr[i, k]
            s[9 * i:9 * i + 9, 9 * j:9 * j + 9] += model["block"][0, k]
    return s


def configurations(model, frozen_species: int | None):
    """(species per site, probability) of every configuration; site 0 fixed to frozen_species if given."""
    cls = model["cls"]
    choices = []
    for i, kap in enumerate(cls):
        if i == 0 and frozen_species is not None:
            choices.append([(frozen_species, 1.0)])
        else:
            choices.append(list(zip(model["spc"][kap], model["conc"][kap])))
    for combo in itertools.product(*choices):
        sp = np.array([c[0] for c in combo])
        p = float(np.prod([c[1] for c in combo]))
        yield sp, p


def dense_h(model, par, species: np.ndarray, with_e: bool = True) -> np.ndarray:
    """H = E_nu + h - h o h, h = c + d S d, for one configuration (species per site)."""
    s = dense_s(model)
    c = par["c"][species].reshape(-1)
    d = par["d"][species].reshape(-1)
    o = par["o"][species].reshape(-1)
    e = par["e"][species].reshape(-1)
    h = d[:, None] * s * d[None, :] + np.diag(c)
    hm = h - h @ np.diag(o) @ h
    if with_e:
        hm = hm + np.diag(e)
    return hm


def exact_average_g(model, par, z: complex, iv: int, frozen_species: int | None, ep: bool = False) -> complex:
    """<<[(z - K)^-1]_{iv, iv}>> over the configurations; K = H (with E_nu) or, ep=True, K = H(e = 0) + diag(z pa + pb)."""
    tot = 0.0
    for sp, p in configurations(model, frozen_species):
        if ep:
            k = dense_h(model, par, sp, with_e=False)
            x = z * par["pa"][sp].reshape(-1) + par["pb"][sp].reshape(-1)
            k = k + np.diag(x)
        else:
            k = dense_h(model, par, sp)
        rhs = np.zeros(k.shape[0], np.complex128)
        rhs[iv] = 1.0
        sol = np.linalg.solve(z * np.eye(k.shape[0]) - k, rhs)
        tot += p * sol[iv]
    return tot


def dense_operator(apply, dim: int) -> np.ndarray:
    """Dense matrix of a linear map x -> apply(x) (columns = images of the unit vectors)."""
    m = np.zeros((dim, dim))
    for i in range(dim):
        u = np.zeros(dim)
        u[i] = 1.0
        m[:, i] = apply(u)
    return m


def dense_asa(model, pp, species: np.ndarray):
    """h, O and H of the ASA secular problem K = E O - H for one configuration, in the form of bands_core (alpha
    representation): h = c + D^1/2 S D^1/2, O = 1 + o h + h o + h p h, H = E_nu + h + E_nu o h + h o E_nu + h (o + p E_nu) h.
    pp: per species and orbital (nsp, 9) arrays c (= C - E_nu), dl (Delta), enu, o, p."""
    s = dense_s(model)
    c, dl, en, ov, pv = (pp[k][species].reshape(-1) for k in ("c", "dl", "enu", "o", "p"))
    sd = np.sqrt(dl)
    h = sd[:, None] * s * sd[None, :] + np.diag(c)
    om = np.eye(h.shape[0]) + ov[:, None] * h + h * ov[None, :] + h @ np.diag(pv) @ h
    hm = np.diag(en) + h + (en * ov)[:, None] * h + h * (en * ov)[None, :] + h @ np.diag(ov + pv * en) @ h
    return h, om, hm


def exact_average_asa(model, pp, z: complex, iv: int, frozen_species: int):
    """Configuration averages of G_11, (h G)_11 and (h G h)_11 with G = (z O - H)^-1, the start site frozen (the brute
    force of test/test_asr_ed3.f90:312-345 restated)."""
    g11 = gh = ghh = 0.0
    for sp, p in configurations(model, frozen_species):
        h, om, hm = dense_asa(model, pp, sp)
        rhs = np.zeros((h.shape[0], 2), np.complex128)
        rhs[iv, 0] = 1.0
        rhs[:, 1] = h[:, iv]
        sol = np.linalg.solve(z * om - hm, rhs)
        g11 += p * sol[iv, 0]
        gh += p * (h[iv, :] @ sol[:, 0])
        ghh += p * (h[:, iv] @ sol[:, 1])
    return g11, gh, ghh


def pair_configurations(model, start_species: int, partner: int, partner_species: int):
    """(species per site, probability) with site 0 and the pair partner frozen (the conditional average of the EPI)."""
    for sp, p in configurations(model, start_species):
        if sp[partner] != partner_species:
            continue
        kap = model["cls"][partner]
        pr = model["conc"][kap][list(model["spc"][kap]).index(partner_species)]
        yield sp, p / pr


def exact_average_pair_blocks(model, sites_par, z: complex, v9, partner: int, partner_species: int, start_species: int):
    """Conditional averages of X0 = v9^T G_00 v9, X1 = v9^T (hG)_00 v9, X2 = v9^T (hGh)_00 v9 with G = (z - H)^-1,
    H = E_nu + h - h o h, h = c + d S d (unscaled, per-site species parameters sites_par[k] = (nsp, 9) arrays for
    k = c, d, o, e; e = E_nu); site 0 carries start_species, the partner partner_species (brute force over all
    configurations of the other sites)."""
    s = dense_s(model)
    x = np.zeros(3, np.complex128)
    v = np.zeros(s.shape[0])
    v[:9] = v9
    for sp, p in pair_configurations(model, start_species, partner, partner_species):
        c, d, o, e = (sites_par[k][sp].reshape(-1) for k in ("c", "d", "o", "e"))
        h = d[:, None] * s * d[None, :] + np.diag(c)
        hm = np.diag(e) + h - h @ np.diag(o) @ h
        hv = h @ v
        sol = np.linalg.solve(z * np.eye(h.shape[0]) - hm, np.stack([v, hv], axis=1))
        x += p * np.array([v @ sol[:, 0], hv @ sol[:, 0], hv @ sol[:, 1]])
    return x


def pair_hamiltonians(model, sites_par, partner: int, partner_species: int, start_species: int):
    """(H, probability) per configuration with site 0 on start_species and the partner frozen: H = E_nu + h - h o h with the
    per-species parameter rows sites_par[k] (k = c, d, o, e)."""
    for sp, p in pair_configurations(model, start_species, partner, partner_species):
        yield dense_h(model, sites_par, sp), p


def fermi(e, mu: float, kt: float):
    t = (np.asarray(e) - mu) / kt
    return np.where(t > 0, np.exp(-np.abs(t)) / (1.0 + np.exp(-np.abs(t))), 1.0 / (1.0 + np.exp(-np.abs(t))))


def exact_pair_domega(model, sites_of, lam: float, partner: int, partner_species: int, mu: float, kt: float,
                      dl: float = 1.0e-5) -> float:
    """dOmega/dlambda = sum_configs p sum_i f(e_i) <psi_i|dH/dlambda|psi_i> (Hellmann-Feynman, brute force), site 0 on
    the path point at lambda (sites_of(lam): parameter rows with the path point as the last species), dH/dlambda by a
    central difference."""
    tot = 0.0
    hp = list(pair_hamiltonians(model, sites_of(lam + dl), partner, partner_species, len(sites_of(lam)["c"]) - 1))
    hm = list(pair_hamiltonians(model, sites_of(lam - dl), partner, partner_species, len(sites_of(lam)["c"]) - 1))
    for (h0, p), (h1, _), (h2, _) in zip(pair_hamiltonians(model, sites_of(lam), partner, partner_species,
                                                           len(sites_of(lam)["c"]) - 1), hp, hm):
        e, v = np.linalg.eigh(h0)
        dh = (h1 - h2) / (2.0 * dl)
        tot += p * float(np.sum(fermi(e, mu, kt) * np.einsum("ji,jk,ki->i", v, dh, v)))
    return tot


def exact_pair_omega(model, sites_par, partner: int, partner_species: int, start_species: int, mu: float, kt: float):
    """Omega = sum_configs p sum_i -kT log(1 + exp(-(e_i - mu) / kT)) (brute force, fixed E_F)."""
    tot = 0.0
    for h, p in pair_hamiltonians(model, sites_par, partner, partner_species, start_species):
        t = -(np.linalg.eigvalsh(h) - mu) / kt
        tot += p * float(np.sum(-kt * np.logaddexp(0.0, t)))
    return tot
