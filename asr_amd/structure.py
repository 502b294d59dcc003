"""Host-side inputs of the augmented-space operator: periodic graphs, screened structure constants, point groups.

These are the fixture generators the chain engine consumes (the Fortran driver builds the same arrays before it calls the
engine).  Each function restates one reference routine; indices are 0-based.

    gauss_legendre          src/asr_harmonics.f90:256-276 (same Newton iteration as src/asr_contour.f90:132-151)
    ylm_real                src/asr_harmonics.f90:177-213
    gaunt_table             src/asr_harmonics.f90:227-254
    bare_s0_block           src/asr_strconst.f90:99-121
    lattice_cluster         src/asr_strconst.f90:53-94
    screened_cluster        src/asr_strconst.f90:126-157
    periodic_graph_shells   src/asr_lattices.f90:49-109
    periodic_graph          src/asr_lattices.f90:14-45, 111-136
    cubic_site_perms        src/asr_symmetry.f90:13-42
    class_ops               src/asr_symmetry.f90:67-92
    orbital_rep             src/asr_symmetry.f90:100-129
    lmto_alloy_blocks       src/asr_lmto_alloy.f90:315-354 (bond blocks only)
    augn_basis, class_mt    src/asr_augcls.f90:73-91, 190-205
    sro_conditional         src/asr_sro.f90:28-80
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

ORB_L = np.array([0, 1, 1, 1, 2, 2, 2, 2, 2], dtype=np.int64)
STANDARD_ALPHA = np.array([0.3485, 0.05303, 0.010714])


def lm_idx(l: int, m: int) -> int:
    return l * l + l + m


def gauss_legendre(n: int):
    x = np.zeros(n)
    w = np.zeros(n)
    for i in range(1, (n + 1) // 2 + 1):
        z = math.cos(math.pi * (i - 0.25) / (n + 0.5))
        pp = 1.0
        for _ in range(100):
            p1, p2 = 1.0, 0.0
            for k in range(1, n + 1):
                p3 = p2
                p2 = p1
                p1 = ((2 * k - 1) * z * p2 - (k - 1) * p3) / k
            pp = n * (z * p1 - p2) / (z * z - 1.0)
            z1 = z
            z = z1 - p1 / pp
            if abs(z - z1) < 1.0e-15:
                break
        x[i - 1] = -z
        x[n - i] = z
        w[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp)
        w[n - i] = w[i - 1]
    return x, w


def _fact_ratio(l: int, m: int) -> float:
    f = 1.0
    for k in range(l - m + 1, l + m + 1):
        f /= float(k)
    return f


def ylm_real(lmax: int, x: float, y: float, z: float) -> np.ndarray:
    """Real orthonormal Y_lm, m = -l..l within a shell (s | y z x | xy yz 3z2-r2 xz x2-y2)."""
    r = math.sqrt(x * x + y * y + z * z)
    ct = z / r
    st = math.sqrt(max(0.0, 1.0 - ct * ct))
    phi = math.atan2(y, x)
    p = np.zeros((lmax + 1, lmax + 1))
    p[0, 0] = 1.0
    for m in range(1, lmax + 1):
        p[m, m] = p[m - 1, m - 1] * (2 * m - 1) * st
    for m in range(0, lmax):
        p[m + 1, m] = ct * (2 * m + 1) * p[m, m]
    for m in range(0, lmax + 1):
        for l in range(m + 2, lmax + 1):
            p[l, m] = (ct * (2 * l - 1) * p[l - 1, m] - (l + m - 1) * p[l - 2, m]) / float(l - m)
    yl = np.zeros((lmax + 1) ** 2)
    for l in range(lmax + 1):
        for m in range(l + 1):
            nrm = math.sqrt((2 * l + 1) / (4.0 * math.pi) * _fact_ratio(l, m))
            if m == 0:
                yl[lm_idx(l, 0)] = nrm * p[l, 0]
            else:
                yl[lm_idx(l, m)] = math.sqrt(2.0) * nrm * p[l, m] * math.cos(m * phi)
                yl[lm_idx(l, -m)] = math.sqrt(2.0) * nrm * p[l, m] * math.sin(m * phi)
    return yl


_GAUNT = {}


def gaunt_table(lmx: int) -> np.ndarray:
    """G(L, L', L'') = int Y_L Y_L' Y_L'' dOmega by Gauss-Legendre x trapezoid quadrature (exact for these degrees)."""
    if lmx in _GAUNT:
        return _GAUNT[lmx]
    n1 = (lmx + 1) ** 2
    n3 = (2 * lmx + 1) ** 2
    g = np.zeros((n1, n1, n3))
    nth = 3 * lmx + 3
    nph = 6 * lmx + 6
    xg, wg = gauss_legendre(nth)
    for it in range(nth):
        ct = xg[it]
        st = math.sqrt(1.0 - ct * ct)
        for ip in range(nph):
            ph = 2.0 * math.pi * ip / nph
            w = wg[it] * 2.0 * math.pi / nph
            yv = ylm_real(2 * lmx, st * math.cos(ph), st * math.sin(ph), ct)
            g += w * np.einsum("i,j,k->ijk", yv[:n1], yv[:n1], yv[:n3])
    g[np.abs(g) < 1.0e-13] = 0.0
    _GAUNT[lmx] = g
    return g


def avg_ws_radius(kind: str) -> float:
    per = {"sc": 4.0, "bcc": 8.0, "fcc": 16.0}[kind]
    return (3.0 / (per * math.pi)) ** (1.0 / 3.0)


def hardcore_alpha(sigma: float) -> np.ndarray:
    return np.array([sigma ** (2 * l + 1) / (2.0 * (2 * l + 1)) for l in range(4)])


def bare_s0_block(d, w: float) -> np.ndarray:
    """Bare canonical 9x9 block S0_{LL'}(d) (rows L, columns L'); d in units of alat, w the average WS radius."""
    gt = gaunt_table(2)
    d = np.asarray(d, dtype=float)
    r = math.sqrt(float(d @ d))
    yv = ylm_real(4, d[0], d[1], d[2])
    df = [1.0, 1.0, 3.0, 15.0, 105.0]
    blk = np.zeros((9, 9))
    for j in range(9):
        lj = int(ORB_L[j])
        for i in range(9):
            li = int(ORB_L[i])
            ll = li + lj
            s = 0.0
            for mm in range(-ll, ll + 1):
                s += gt[i, j, lm_idx(ll, mm)] * yv[lm_idx(ll, mm)]
            fac = -8.0 * math.pi * (-1.0) ** lj * df[ll] / (df[li] * df[lj]) * (w / r) ** (ll + 1)
            blk[i, j] = fac * s
    return blk


def _on_lat(kind: str, i: int, j: int, k: int) -> bool:
    if kind == "sc":
        return i % 2 == 0 and j % 2 == 0 and k % 2 == 0
    if kind == "bcc":
        return i % 2 == j % 2 and j % 2 == k % 2
    if kind == "fcc":
        return (i + j + k) % 2 == 0
    return False


def lattice_cluster(kind: str, rmax: float) -> np.ndarray:
    """Lattice vectors (alat) within rmax, origin first, then by distance (stable insertion sort); shape (n, 3)."""
    nmax = math.ceil(2.0 * rmax) + 1
    pts = []
    for k in range(-nmax, nmax + 1):
        for j in range(-nmax, nmax + 1):
            for i in range(-nmax, nmax + 1):
                ip = (i, j, k)
                if kind == "sc" and any(v % 2 != 0 for v in ip):
                    continue
                if kind == "bcc" and any(v % 2 != i % 2 for v in ip):
                    continue
                if kind == "fcc" and sum(ip) % 2 != 0:
                    continue
                p = 0.5 * np.array(ip, dtype=float)
                if math.sqrt(float(p @ p)) > rmax + 1.0e-9:
                    continue
                pts.append(p)
    v = [p.copy() for p in pts]
    for i in range(1, len(v)):
        t = v[i]
        di = float(t @ t)
        j = i - 1
        while j >= 0:
            if float(v[j] @ v[j]) <= di + 1.0e-12:
                break
            v[j + 1] = v[j]
            j -= 1
        v[j + 1] = t
    return np.array(v)


def screened_cluster(vec: np.ndarray, w: float, alpha) -> np.ndarray:
    """Screened S^alpha_{R0} blocks (n, 9, 9) from (1 - S0 alpha) S^alpha = S0 on the cluster (origin = site 0)."""
    n = vec.shape[0]
    nn = 9 * n
    s0 = np.zeros((nn, nn))
    for j in range(n):
        for i in range(n):
            if i == j:
                continue
            s0[9 * i:9 * i + 9, 9 * j:9 * j + 9] = bare_s0_block(vec[i] - vec[j], w)
    ad = np.asarray(alpha, dtype=float)[ORB_L]
    m = -s0 * np.tile(ad, n)[None, :]
    m[np.diag_indices(nn)] += 1.0
    x = np.linalg.solve(m, s0[:, 0:9])
    return x.reshape(n, 9, 9)


@dataclass
class Graph:
    """Symmetric neighbour graph: nbr[site, slot] (-1: no neighbour on the slot), coord (n, 3), bvec (nslot, 3)."""
    n: int
    nbr: np.ndarray
    coord: np.ndarray
    bvec: np.ndarray | None = None

    @property
    def nbr_ptr(self) -> np.ndarray:
        return np.arange(self.n + 1, dtype=np.int32) * self.nbr.shape[1]


def _site_index(kind: str, l: int):
    m = 2 * l
    idx = -np.ones((m, m, m), dtype=np.int64)
    coord = []
    for k in range(m):
        for j in range(m):
            for i in range(m):
                if _on_lat(kind, i, j, k):
                    idx[i, j, k] = len(coord)
                    coord.append((i, j, k))
    return idx, np.array(coord, dtype=np.int64)


def periodic_graph_shells(kind: str, l: int, r2max: int) -> Graph:
    """Periodic graph with every neighbour within r2max (half-lattice units squared); slot k = bond vector bvec[k]."""
    m = 2 * l
    idx, coord = _site_index(kind, l)
    rmax = int(math.sqrt(float(r2max))) + 1
    tmp = []
    for k in range(-rmax, rmax + 1):
        for j in range(-rmax, rmax + 1):
            for i in range(-rmax, rmax + 1):
                if i == 0 and j == 0 and k == 0:
                    continue
                if kind != "sc" and not _on_lat(kind, i % 2, j % 2, k % 2):
                    continue
                if kind == "sc" and any(v % 2 != 0 for v in (i, j, k)):
                    continue
                if i * i + j * j + k * k > r2max:
                    continue
                tmp.append((i, j, k))
    key = [(a * a + b * b + c * c) * 100000 + (a + 50) * 10000 + (b + 50) * 100 + (c + 50) for a, b, c in tmp]
    order = sorted(range(len(tmp)), key=lambda q: key[q])  # stable, as the insertion sort
    bvec = np.array([tmp[q] for q in order], dtype=np.int64)
    q = (coord[:, None, :] + bvec[None, :, :]) % m
    nbr = idx[q[..., 0], q[..., 1], q[..., 2]].astype(np.int32)
    return Graph(n=coord.shape[0], nbr=nbr, coord=coord, bvec=bvec)


def periodic_graph(kind: str, l: int) -> Graph:
    """Nearest-neighbour periodic graph (sc: 6, bcc: 8, fcc: 12 neighbours)."""
    m = 2 * l
    idx, coord = _site_index(kind, l)
    if kind == "sc":
        vec = []
        for i in range(3):
            for j in (-1, 1):
                v = [0, 0, 0]
                v[i] = 2 * j
                vec.append(v)
    elif kind == "bcc":
        vec = [[i, j, k] for i in (-1, 1) for j in (-1, 1) for k in (-1, 1)]
    else:
        vec = []
        for i in (-1, 1):
            for j in (-1, 1):
                vec += [[i, j, 0], [i, 0, j], [0, i, j]]
    vec = np.array(vec, dtype=np.int64)
    q = (coord[:, None, :] + vec[None, :, :]) % m
    nbr = idx[q[..., 0], q[..., 1], q[..., 2]].astype(np.int32)
    return Graph(n=coord.shape[0], nbr=nbr, coord=coord, bvec=vec)


_AX = [(1, 2, 3), (1, 3, 2), (2, 1, 3), (2, 3, 1), (3, 1, 2), (3, 2, 1)]


def cubic_site_perms(coord: np.ndarray, m: int):
    """perm (48, n): image of every site under the 48 signed axis permutations about the origin; mats (48, 3, 3)."""
    n = coord.shape[0]
    idx = -np.ones((m, m, m), dtype=np.int64)
    idx[coord[:, 0], coord[:, 1], coord[:, 2]] = np.arange(n)
    perm = np.zeros((48, n), dtype=np.int32)
    mats = np.zeros((48, 3, 3), dtype=np.int64)
    iop = 0
    for pp in _AX:
        pp0 = [p - 1 for p in pp]
        for s in range(8):
            sg = [1 - 2 * (s & 1), 1 - 2 * ((s >> 1) & 1), 1 - 2 * ((s >> 2) & 1)]
            for i in range(3):
                mats[iop, i, pp0[i]] = sg[i]
            q = np.stack([(sg[i] * coord[:, pp0[i]]) % m for i in range(3)], axis=1)
            img = idx[q[:, 0], q[:, 1], q[:, 2]]
            if np.any(img < 0):
                raise ValueError("cubic_site_perms: image is not a lattice site")
            perm[iop] = img
            iop += 1
    return perm, mats


def class_ops(coord: np.ndarray, m: int, cls: np.ndarray, group: str = "d4h"):
    """Class-preserving operations about the origin: 'd4h' (keeps the z axis, the reference's class_ops), 'd2h' (diagonal
    sign operations: every one of the 9 real harmonics is one-dimensional), 'identity'.  Operation 0 is the identity."""
    p48, m48 = cubic_site_perms(coord, m)
    keep = np.array([np.all(cls[p48[k]] == cls) for k in range(48)])
    if group == "identity":
        keep[:] = False
        keep[0] = True
    else:
        keep &= np.abs(m48[:, 2, 2]) == 1
        if group == "d2h":
            keep &= np.array([np.count_nonzero(m48[k] - np.diag(np.diag(m48[k]))) == 0 for k in range(48)])
    if not keep[0]:
        raise ValueError("class_ops: the identity does not preserve the classes")
    return p48[keep].copy(), m48[keep].copy()


def orbital_rep(mats: np.ndarray) -> np.ndarray:
    """d (nop, 9, 9): Y_k(M p) = sum_j d[k, j] Y_j(p), fitted on 24 fixed points; near-integers snapped (monomial ops exact)."""
    np_ = 24
    wq = np.array([0.414213562373095, 0.732050807568877, 0.236067977499790])
    pts = np.zeros((np_, 3))
    a = np.zeros((np_, 9))
    for ip in range(np_):
        p = (ip + 1) * wq
        p = p - np.trunc(p) - 0.5
        pts[ip] = p
        a[ip] = ylm_real(2, p[0], p[1], p[2])[:9]
    ata = a.T @ a
    d = np.zeros((mats.shape[0], 9, 9))
    for iop in range(mats.shape[0]):
        b = np.zeros((np_, 9))
        for ip in range(np_):
            mp = mats[iop].astype(float) @ pts[ip]
            b[ip] = ylm_real(2, mp[0], mp[1], mp[2])[:9]
        x = np.linalg.solve(ata, a.T @ b)
        dd = x.T.copy()
        near = np.abs(dd - np.rint(dd)) < 1.0e-9
        dd[near] = np.rint(dd[near])
        d[iop] = dd
    return d


def lmto_alloy_blocks(kind: str, rcl: float, alpha, rinv: float | None = None):
    """Bond blocks of build_lmto_alloy: (graph slots of r2max = int((2 rcl)^2), block (nslot, 9, 9), block0 (9, 9)).
    block[k] = S_{R, R + bvec(k)} = (S(-v) + S(v)^T) / 2, block0 = symmetrised on-site S^alpha."""
    w = avg_ws_radius(kind)
    vec = lattice_cluster(kind, max(rcl, rinv) if rinv is not None else rcl)
    blk = screened_cluster(vec, w, alpha)
    r2max = int((2.0 * rcl) ** 2 + 1.0e-9)
    return vec, blk, r2max, w


def bond_blocks_for(bvec: np.ndarray, vec: np.ndarray, blk: np.ndarray):
    def find(u):
        for i in range(vec.shape[0]):
            if np.max(np.abs(vec[i] - u)) < 1.0e-7:
                return i
        raise ValueError("bond vector missing from the screened cluster")

    ns = bvec.shape[0]
    block = np.zeros((ns, 9, 9))
    for k in range(ns):
        u = 0.5 * bvec[k].astype(float)
        ia = find(-u)
        ib = find(u)
        block[k] = 0.5 * (blk[ia] + blk[ib].T)
    block0 = 0.5 * (blk[0] + blk[0].T)
    return block, block0


def augn_basis(conc) -> np.ndarray:
    """phi (ns, ns): row l is phi_l on the species, orthonormal in <f, g> = sum_s x_s f(s) g(s), phi_0 = 1."""
    conc = np.asarray(conc, dtype=float)
    ns = conc.size
    phi = np.zeros((ns, ns))
    phi[0, :] = 1.0
    for l in range(1, ns):
        phi[l, :] = 0.0
        phi[l, l - 1] = 1.0
        for _ in range(2):
            for m in range(l):
                ip = np.sum(conc * phi[m] * phi[l])
                phi[l] = phi[l] - ip * phi[m]
        phi[l] = phi[l] / math.sqrt(np.sum(conc * phi[l] ** 2))
    return phi


def class_mt(conc, nsmax: int) -> np.ndarray:
    """mt[l, k, s] = x_s phi_l(s) phi_k(s) on the species of non-zero weight (the others stay zero)."""
    conc = np.asarray(conc, dtype=float)
    mt = np.zeros((nsmax, nsmax, nsmax))
    sup = np.nonzero(conc > 0.0)[0]
    n = sup.size
    phi = augn_basis(conc[sup])
    for l in range(n):
        for k in range(n):
            for j in range(n):
                mt[l, k, sup[j]] = conc[sup[j]] * phi[l, j] * phi[k, j]
    return mt


def sro_conditional(x, y, alpha, same: bool = False):
    """Conditional concentrations cond[c, s] = P(species s on a first-shell neighbour | species c at the centre)
    = y_s (1 - alpha_cs) for the Warren-Cowley table alpha (n x m); one sublattice (same): alpha symmetric with the diagonal
    derived, alpha_cc = -sum_s x_s alpha_cs / x_c; two sublattices: both marginal conditions must hold."""
    x = np.asarray(x, dtype=float)
    y = np.asarray(y, dtype=float)
    alpha = np.array(alpha, dtype=float)
    n, m = x.size, y.size
    if alpha.shape != (n, m):
        raise ValueError("sro: alpha must be n(centre) x n(neighbour)")
    if np.any(x <= 0) or np.any(y <= 0) or abs(x.sum() - 1) > 1e-8 or abs(y.sum() - 1) > 1e-8:
        raise ValueError("sro: concentrations must be > 0 and sum to 1")
    if same:
        if n != m or np.max(np.abs(x - y)) > 1e-10:
            raise ValueError("sro: one sublattice needs x = y")
        if np.any(np.abs(np.diag(alpha)) > 1e-10) or np.max(np.abs(alpha - alpha.T)) > 1e-10:
            raise ValueError("sro: alpha must be symmetric with a zero diagonal on one sublattice")
        for i in range(n):
            alpha[i, i] = -np.sum(x * alpha[i, :]) / x[i]
    else:
        if np.max(np.abs(alpha @ y)) > 1e-8 or np.max(np.abs(x @ alpha)) > 1e-8:
            raise ValueError("sro: the marginal conditions sum_s y_s alpha_cs = 0 = sum_c x_c alpha_cs do not hold")
    cond = y[None, :] * (1.0 - alpha)
    if np.any(cond < -1e-10):
        raise ValueError("sro: a conditional probability is negative")
    cond[np.abs(cond) <= 1e-10] = 0.0
    if np.max(np.abs(cond.sum(axis=1) - 1.0)) > 1e-8:
        raise ValueError("sro: internal error, conditional sum /= 1")
    return cond, alpha
