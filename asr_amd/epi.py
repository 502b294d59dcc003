"""Host side of the effective pair interactions on the chain engine (src/asr_epi.f90).

Thermodynamic integration over the species of site 0: its H^(2) parameters go from species c0 to c on a straight path
(epi_path); at each path point the central blocks X0 = G_00, X1 = (hG)_00, X2 = (hGh)_00 of every symmetry-adapted start
vector come from two complex-symmetric chains (asr_gpu_epi_blocks), and
    dOmega/dlambda = -(1/pi) Im sum_j w_j sum_L epi_integrand(pt, L, z_j, X(z_j)),
U(c, s) = int_0^1 dlambda [dOmega/dlambda (c0 -> c | s) - dOmega/dlambda (c0 -> c | s0)].

This module holds the species parameters (epi_species without the representation change, epi_path, epi_setup
src/asr_epi.f90:124-178), the trace formula (epi_integrand :182-189), dOmega/dlambda over a contour (epi_domega), the pair
table entry by the path quadrature (epi_pair_u), the symmetry-adapted start vectors of an abelian pair group (epi_adapted
:290-328) and the pair-table algebra (epi_exchange :257-266).  The chains run on the GPU
(asr_amd.engine.GpuOperator.epi_blocks).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .augcls import Member

ORB_L = np.array([0, 1, 1, 1, 2, 2, 2, 2, 2])


@dataclass
class EpiPar:
    """One species in the screened representation, per orbital: E_nu, C - E_nu, |sqrt(Delta)|, o (epi_par_t)."""
    e: np.ndarray
    c: np.ndarray
    d: np.ndarray
    o: np.ndarray


@dataclass
class EpiPath:
    """Site 0 at lambda on the path a -> b: the parameters and their derivatives (d/dlambda of E_nu, c, ln d, o)."""
    e: np.ndarray
    c: np.ndarray
    d: np.ndarray
    o: np.ndarray
    de: np.ndarray
    dc: np.ndarray
    dk: np.ndarray
    dox: np.ndarray


def epi_species(pp_l: dict) -> EpiPar:
    """Per-orbital parameters from per-l values (3,) in the alpha representation: enu, c (= C - E_nu), sqrdel, o."""
    return EpiPar(e=np.asarray(pp_l["enu"], float)[ORB_L], c=np.asarray(pp_l["c"], float)[ORB_L],
                  d=np.abs(np.asarray(pp_l["sqrdel"], float))[ORB_L], o=np.asarray(pp_l["o"], float)[ORB_L])


def epi_path(a: EpiPar, b: EpiPar, lam: float) -> EpiPath:
    """Straight path from species a (lambda = 0) to b (lambda = 1) (epi_path, src/asr_epi.f90:139-148)."""
    d = a.d + lam * (b.d - a.d)
    return EpiPath(e=a.e + lam * (b.e - a.e), c=a.c + lam * (b.c - a.c), d=d, o=a.o + lam * (b.o - a.o),
                   de=b.e - a.e, dc=b.c - a.c, dk=(b.d - a.d) / d, dox=b.o - a.o)


def epi_setup(par: list[EpiPar], pt: EpiPath):
    """Operator parameters for species 0..ns-1 and the path point (species ns) in the scaling of route edc with the Delta
    of the path point (epi_setup, src/asr_epi.f90:153-178): c = (D_c / D_s) c_s, d = D_c^1/2, o = o_s D_s / D_c, e = 0,
    and the energy-dependent diagonal e_s(z) = z - D_c (z - E_nu,s) / D_s = z pa + pb.  Returns c, d, o, e, pa, pb as
    (ns + 1, 9) arrays."""
    ns = len(par)
    dc = pt.d ** 2
    c = np.zeros((ns + 1, 9))
    d = np.zeros((ns + 1, 9))
    o = np.zeros((ns + 1, 9))
    dl = np.zeros((ns + 1, 9))
    en = np.zeros((ns + 1, 9))
    for s, p in enumerate(par):
        ds = p.d ** 2
        c[s] = dc / ds * p.c
        d[s] = pt.d
        o[s] = p.o * ds / dc
        dl[s] = ds
        en[s] = p.e
    c[ns], d[ns], o[ns], dl[ns], en[ns] = pt.c, pt.d, pt.o, dc, pt.e
    pa = 1.0 - dc[None, :] / dl
    pb = dc[None, :] * en / dl
    return c, d, o, np.zeros((ns + 1, 9)), pa, pb


def epi_member(par: list[EpiPar], pt: EpiPath, l: int, chi: np.ndarray, partner_species: int, z: complex) -> Member:
    """One EPI chain block: site 0 frozen to the path point (species ns), the partner frozen to partner_species, the
    characters chi of the start vector, energy z; l is an orbital index of the start vector's l shell."""
    c, d, o, e, pa, pb = epi_setup(par, pt)
    return Member(l0=int(l), c=c, d=d, o=o, e=e, frozen=len(par), frozen2=int(partner_species), pa=pa, pb=pb,
                  z=complex(z), chi=np.asarray(chi, float))


def epi_integrand(pt: EpiPath, l: int, z: complex, x) -> complex:
    """[G dH/dlambda]_LL for one start vector of orbital character l from its blocks x = (X0, X1, X2) at z, without the
    constant -2 kappa (epi_integrand, src/asr_epi.f90:182-189)."""
    return ((pt.de[l] + pt.dc[l]) * x[0] - 2.0 * pt.dc[l] * pt.o[l] * x[1] - pt.dox[l] * x[2]
            + 2.0 * pt.dk[l] * ((z - pt.e[l] - pt.c[l]) * x[0] + 2.0 * pt.c[l] * pt.o[l] * x[1] - pt.o[l] * x[2]))


def epi_domega(G, par: list[EpiPar], pt: EpiPath, partner_species: int, zs, ws, vec: np.ndarray, chi: np.ndarray,
               lo: np.ndarray, nlev: int, mode: int = 0, max_blocks: int = 56) -> float:
    """dOmega/dlambda at the path point pt with the pair partner frozen to partner_species, without the constant -2 kappa
    term of epi_integrand (the same for every partner species, so it cancels in the pair table):
        -(1/pi) Im sum_j ws_j sum_a epi_integrand(pt, lo_a, zs_j, X_a(zs_j))
    over the contour (zs, ws) and the nine adapted start vectors (vec[:, a], characters chi[:, a], shell lo[a] of
    epi_adapted), the blocks X_a from the GPU (GpuOperator.epi_blocks of the pair operator G; calls of at most max_blocks
    blocks) (the dOmega/dlambda of the EPI driver, src/asr_epi.f90:241 with epi_integrand :182-189)."""
    zs = np.asarray(zs, np.complex128).ravel()
    ws = np.asarray(ws, np.complex128).ravel()
    tasks = [(a, j) for j in range(zs.size) for a in range(9)]
    tot = 0.0 + 0.0j
    for q0 in range(0, len(tasks), max_blocks):
        chunk = tasks[q0:q0 + max_blocks]
        members = [epi_member(par, pt, int(lo[a]), chi[:, a], partner_species, zs[j]) for a, j in chunk]
        v9 = np.stack([vec[:, a] for a, _ in chunk])
        x, *_ = G.epi_blocks(members, v9, nlev, mode=mode)
        for q, (a, j) in enumerate(chunk):
            tot += ws[j] * epi_integrand(pt, int(lo[a]), zs[j], x[q])
    return float(-tot.imag / np.pi)


def epi_pair_u(G, par: list[EpiPar], c0: int, c: int, s: int, s0: int, zs, ws, vec, chi, lo, nlev: int, nlam: int = 2,
               mode: int = 0, max_blocks: int = 56) -> float:
    """U(c, s) = int_0^1 dlambda [dOmega/dlambda (c0 -> c | s) - dOmega/dlambda (c0 -> c | s0)] by nlam Gauss-Legendre
    points on the straight path of site 0 from species c0 to c (the pair table entry against the first species c0, s0,
    src/asr_epi.f90:4-24)."""
    from .host import gauss_legendre
    x, w = gauss_legendre(nlam)
    u = 0.0
    for xk, wk in zip(x, w):
        pt = epi_path(par[c0], par[c], 0.5 * (1.0 + xk))
        u += 0.5 * wk * (epi_domega(G, par, pt, s, zs, ws, vec, chi, lo, nlev, mode, max_blocks)
                         - epi_domega(G, par, pt, s0, zs, ws, vec, chi, lo, nlev, mode, max_blocks))
    return u


def epi_exchange(u: np.ndarray) -> np.ndarray:
    """V(p, q) = U(p, p) + U(q, q) - U(p, q) - U(q, p) of one sublattice (epi_exchange, src/asr_epi.f90:257-266)."""
    dg = np.diag(u)
    return dg[:, None] + dg[None, :] - u - u.T


def epi_adapted(dmat: np.ndarray):
    """Symmetry-adapted start vectors of an abelian group of involutions with orbital matrices dmat (nop, 9, 9)
    (epi_adapted, src/asr_epi.f90:290-328): vec[:, a] orthonormal, each inside one l shell, D(g) vec = chi[g, a] vec with
    chi = +-1; lo[a] the first orbital index of that shell.  With the identity alone the vectors are the harmonics."""
    nop = dmat.shape[0]
    for g in range(nop):
        if np.max(np.abs(dmat[g] - dmat[g].T)) > 1.0e-9:
            raise ValueError("epi_adapted: an operation is not an involution (the group must be abelian of order-2 elements)")
    vec = np.zeros((9, 9))
    lo = np.zeros(9, np.int32)
    for i1, i2 in ((0, 1), (1, 4), (4, 9)):
        # a generic combination of the group's matrices has one eigenvalue per character pattern
        m = sum(np.sqrt(g + 1 + 0.37) * dmat[g, i1:i2, i1:i2] for g in range(nop))
        m = 0.5 * (m + m.T)
        _, v = np.linalg.eigh(m)
        vec[i1:i2, i1:i2] = v
        lo[i1:i2] = i1
    chi = np.zeros((nop, 9))
    for a in range(9):
        v = vec[:, a]
        v[np.abs(v) < 1.0e-12] = 0.0
        k = int(np.argmax(np.abs(v)))
        if v[k] < 0.0:
            v = -v
        v = v / np.linalg.norm(v)
        vec[:, a] = v
        for g in range(nop):
            w = dmat[g] @ v
            x = v @ w
            if abs(abs(x) - 1.0) > 1.0e-8 or np.max(np.abs(w - x * v)) > 1.0e-8:
                raise ValueError("epi_adapted: a start vector is not a common eigenvector of the group")
            chi[g, a] = np.sign(x)
    return vec, chi, lo
