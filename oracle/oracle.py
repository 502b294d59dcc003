"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg).

ctypes bindings of liborc.so (the C restatement in asr_oracle.c) plus small numpy restatements of the reference routines
that call LAPACK on the chain matrices:
    np_cf_species       src/asr_chain.f90:67-167          (zgeev -> numpy.linalg.eigvals, the same LAPACK routine)
    np_cf_closed        src/asr_ed3.f90:214-230           (the fraction of a closed chain, no tail)
    OracleOp.ed3_blocks src/asr_ed3.f90:247-278           (two chains -> the central blocks bk(0:3))
    np_bp_asymptote     src/reference/asr_terminator.f90:39-79  (dsterf -> scipy.linalg.eigvalsh_tridiagonal)
    np_cf_green         src/reference/asr_greens.f90:207-217
    np_sqrt_tail        src/reference/asr_greens.f90:139-149
    np_kpm_dos          src/reference/asr_kpm.f90:68-87
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    subprocess.check_call(["make", "-s", "-C", _HERE])
    if os.path.isdir("/root/reference/src") and os.path.exists("/opt/rocm/bin/amdflang"):
        # the reference's chain modules compile-only and the Fortran binding against them (oracle/_ref/fmod)
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])
    return os.path.join(_HERE, "liborc.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        L.orc_apply.argtypes = [P, P, P, P]
        L.orc_rand_c.argtypes = [P, P, P, P, P]
        L.orc_project_c.argtypes = [P, P, P]
        L.orc_lanczos.argtypes = [P, P, C.c_int, C.c_int, C.c_double, P, P]
        L.orc_lanczos_ro.argtypes = [P, P, C.c_int, C.c_int, C.c_double, C.c_int, P, P]
        L.orc_edc_chain.argtypes = [P, P, P, C.c_int, C.c_int, C.c_int, P, P]
        L.orc_ed3_apply.argtypes = [P, P, P, P, P, P, P]
        L.orc_ed3_start.argtypes = [P, P, C.c_int, C.c_double, C.c_int, P, P, P, P]
        L.orc_ed3_chain.argtypes = [P, P, P, P, P, C.c_double, C.c_double, P, P, C.c_int, C.c_int, P, P, P, P]
        L.orc_epi_start.argtypes = [P, P, P, C.c_int, P, P, P]
        L.orc_epi_chain.argtypes = [P, P, P, P, P, C.c_int, C.c_int, P, P, P, P]
        L.orc_edc_chain_proj.argtypes = [P, P, P, C.c_int, C.c_int, C.c_int, C.c_int, P, P, P, P, P, P]
        L.orc_lanczos_batch.argtypes = [P, P, P, P, C.c_int, C.c_int, P, P, P]
        L.orc_edc_batch.argtypes = [P, P, P, P, P, C.c_int, C.c_int, C.c_int, P, P, P]
        L.orc_real_apply.argtypes = [P, P, P, P]
        L.orc_kpm_moments.argtypes = [P, C.c_double, C.c_double, C.c_int, C.c_int, C.c_uint64, P, P]
        L.orc_real_lanczos.argtypes = [P, C.c_int, C.c_uint64, P, P]
        L.orc_rademacher.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_rademacher.restype = C.c_double
        L.orc_threads.argtypes = []
        L.orc_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class _Op(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nn", "nslot", "nop", "nsmax", "ncls", "nsp", "ntype")] + \
               [(n, C.c_void_p) for n in ("hopr", "hopg", "flipr", "flipg", "lst", "ncl", "ntyp", "stab_ptr", "stab", "pi",
                                          "nsc", "spc", "fz", "sq", "sg", "block", "block0", "mt")]


class _Member(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("chi", "c", "d", "o", "e", "mt")] + [("frozen", C.c_int), ("frozen2", C.c_int)]


class _Real(C.Structure):
    _fields_ = [("nsite", C.c_int), ("nslot", C.c_int), ("nsp", C.c_int)] + \
               [(n, C.c_void_p) for n in ("nbr", "species", "c", "d", "o", "e", "block", "block0")]


class OracleOp:
    """The operator tables of an asr_amd.augcls.AugclsOperator, pinned for the C oracle."""

    def __init__(self, op):
        self.keep = []
        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self.keep.append(a)
            return _p(a)
        self.op = op
        self.s = _Op(op.nn, op.nslot, op.nop, op.nsmax, op.ncls, op.nsp, op.ntype,
                     arr(op.hopr, np.int32), arr(op.hopg, np.int32), arr(op.flipr, np.int32), arr(op.flipg, np.int32),
                     arr(op.lst, np.int32), arr(op.ncl, np.int32), arr(op.ntyp, np.int32), arr(op.stab_ptr, np.int32),
                     arr(op.stab if op.stab.size else np.zeros(1, np.int32), np.int32), arr(op.pi, np.int32),
                     arr(op.nsc, np.int32), arr(op.spc, np.int32), arr(op.fz, np.uint8), arr(op.sq, np.float64),
                     arr(op.sg, np.float64), arr(op.block, np.float64), arr(op.block0, np.float64), arr(op.mt, np.float64))

    def member(self, m):
        keep = []
        def arr(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return _p(a)
        chi = self.op.chi_for(m.l0) if getattr(m, "chi", None) is None else np.asarray(m.chi, float)
        s = _Member(arr(chi), arr(m.c), arr(m.d), arr(m.o), arr(m.e),
                    arr(m.mt(self.op)) if m.conc is not None else None, int(m.frozen), int(getattr(m, "frozen2", -1)))
        s._keep = keep
        return s

    def apply(self, m, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        ms = self.member(m)
        lib().orc_apply(C.byref(self.s), C.byref(ms), _p(x), _p(y))
        return y

    def rand_c(self, m, par: np.ndarray, x: np.ndarray) -> np.ndarray:
        par = np.ascontiguousarray(par, dtype=np.complex128)
        x = np.ascontiguousarray(x, dtype=np.complex128)
        w = np.zeros_like(x)
        ms = self.member(m)
        lib().orc_rand_c(C.byref(self.s), C.byref(ms), _p(par), _p(x), _p(w))
        return w

    def project_c(self, m, x: np.ndarray) -> np.ndarray:
        x = np.array(x, dtype=np.complex128)
        ms = self.member(m)
        lib().orc_project_c(C.byref(self.s), C.byref(ms), _p(x))
        return x

    def lanczos(self, m, iv: int, nmax: int, tiny: float = 1.0e-12, reorth: bool = False):
        a = np.zeros(nmax)
        b = np.zeros(nmax)
        ms = self.member(m)
        n = lib().orc_lanczos_ro(C.byref(self.s), C.byref(ms), int(iv), int(nmax), float(tiny), int(reorth), _p(a), _p(b))
        return a, b, n

    def edc_chain(self, m, ep: np.ndarray, iv: int, nlev: int, project: bool = True):
        ep = np.ascontiguousarray(ep, dtype=np.complex128)
        a = np.zeros(nlev, np.complex128)
        b = np.zeros(nlev, np.complex128)
        ms = self.member(m)
        mm = lib().orc_edc_chain(C.byref(self.s), C.byref(ms), _p(ep), int(iv), int(nlev), int(project), _p(a), _p(b))
        return a, b, mm

    def ed3_apply(self, m, et, uu, qt, x: np.ndarray) -> np.ndarray:
        """K' x (ed3_kmul) with the species coefficients et, uu, qt (nsp, 9) at the chain's energy."""
        cs = [np.ascontiguousarray(v, dtype=np.complex128) for v in (et, uu, qt)]
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros_like(x)
        ms = self.member(m)
        lib().orc_ed3_apply(C.byref(self.s), C.byref(ms), _p(cs[0]), _p(cs[1]), _p(cs[2]), _p(x), _p(y))
        return y

    def ed3_start(self, m, iv: int, oc: float, project: bool = True):
        n = self.op.dim
        vv = np.zeros(n, np.complex128)
        av = np.zeros(n, np.complex128)
        nv = C.c_double()
        na = C.c_double()
        ms = self.member(m)
        st = lib().orc_ed3_start(C.byref(self.s), C.byref(ms), int(iv), float(oc), int(project), _p(vv), _p(av),
                                 C.byref(nv), C.byref(na))
        if st:
            raise ValueError("ed3_start: zero start vector")
        return vv, av, nv.value, na.value

    def ed3_chain(self, m, et, uu, qt, z: complex, u0, tv, nlev: int, project: bool = True):
        cs = [np.ascontiguousarray(v, dtype=np.complex128) for v in (et, uu, qt)]
        u0 = np.ascontiguousarray(u0, dtype=np.complex128)
        tv = np.ascontiguousarray(tv, dtype=np.complex128)
        a = np.zeros(nlev, np.complex128)
        b = np.zeros(nlev, np.complex128)
        c = np.zeros(nlev, np.complex128)
        closed = C.c_int()
        ms = self.member(m)
        mm = lib().orc_ed3_chain(C.byref(self.s), C.byref(ms), _p(cs[0]), _p(cs[1]), _p(cs[2]), complex(z).real,
                                 complex(z).imag, _p(u0), _p(tv), int(nlev), int(project), _p(a), _p(b), _p(c),
                                 C.byref(closed))
        return a, b, c, mm, bool(closed.value)

    def ed3_blocks(self, m, et, uu, qt, oc: float, z: complex, iv: int, nlev: int, project: bool = True,
                   tail: str = "hull"):
        """ed3_blocks: the chain from |a> / na projected on |v>, the chain from |v> / nv projected on |a>;
        bk = (G_11, (hG)_11, (hGh)_11, q1a - q1b).  Returns bk and the two chains' (a, b, c, m, closed)."""
        vv, av, nv, na = self.ed3_start(m, iv, oc, project)
        chains = []
        res = []
        for u0, tv, nrm in ((av / na, vv, na), (vv / nv, av, nv)):
            a, b, c, mm, closed = self.ed3_chain(m, et, uu, qt, z, u0, tv, nlev, project)
            if closed:
                g, f = np_cf_closed(a, b, c, mm, z)
            else:
                g, f = np_cf_species(a, b, c, mm, z, tail)
            chains.append((a, b, c, mm, closed))
            res.append((nrm ** 2 * g, nrm * f))
        (q0, q1a), (q2, q1b) = res
        fm = 0.5 * (q1a + q1b)
        bk = np.array([q0 - 2.0 * oc * fm + oc ** 2 * q2, fm - oc * q2, q2, q1a - q1b], np.complex128)
        return bk, chains, (na, nv)

    def epi_blocks(self, m, ep, v9, z: complex, nlev: int, project: bool = True, tail: str = "hull"):
        """epi_blocks (src/asr_epi.f90:628-647): chain from s0 = v9 projected on |v> = h'|s0>, chain from |v> / nv projected
        on s0; x = (X0, X1, X2, X1a - X1b).  A chain that closed gets no tail.  Returns x and the chains' (a, b, c, m,
        closed)."""
        n = self.op.dim
        ms = self.member(m)
        v9 = np.ascontiguousarray(v9, dtype=np.float64)
        s0 = np.zeros(n, np.complex128)
        vv = np.zeros(n, np.complex128)
        nv = C.c_double()
        if lib().orc_epi_start(C.byref(self.s), C.byref(ms), _p(v9), int(project), _p(s0), _p(vv), C.byref(nv)):
            raise ValueError("epi_start: zero start vector")
        nv = nv.value
        ep = np.ascontiguousarray(ep, dtype=np.complex128)
        res, chains = [], []
        for u0, tv in ((s0, vv), (vv / nv, s0)):
            a = np.zeros(nlev, np.complex128)
            b = np.zeros(nlev, np.complex128)
            c = np.zeros(nlev, np.complex128)
            closed = C.c_int()
            mm = lib().orc_epi_chain(C.byref(self.s), C.byref(ms), _p(ep), _p(np.ascontiguousarray(u0)),
                                     _p(np.ascontiguousarray(tv)), int(nlev), int(project), _p(a), _p(b), _p(c),
                                     C.byref(closed))
            g, f = np_cf_species(a, b, c, mm, z, "none" if closed.value else tail)
            chains.append((a, b, c, mm, bool(closed.value)))
            res.append((g, f))
        (ga, fa), (gv, fv) = res
        fb = nv * fv
        x = np.array([ga, 0.5 * (fa + fb), nv ** 2 * gv, fa - fb], np.complex128)
        return x, chains

    def edc_green_proj(self, m, ep, iv: int, nlev: int, tnode, tgop, tw, z: complex, project: bool = True,
                       tail: str = "hull"):
        """edc_green_proj (src/asr_edc.f90:174-232): g and the projections gp[t, l] = sum_k cpr[k, t, l] pv_k of the chain
        terminated at m, gp4 the same at max(1, m - 4).  tnode, tgop: (nt, nsmax); tw: (nt, nsmax, 9) dressing weights."""
        ep = np.ascontiguousarray(ep, dtype=np.complex128)
        tn = np.ascontiguousarray(tnode, dtype=np.int32)
        tg = np.ascontiguousarray(tgop, dtype=np.int32)
        twa = np.ascontiguousarray(tw, dtype=np.float64)
        nt = tn.shape[0]
        a = np.zeros(nlev, np.complex128)
        b = np.zeros(nlev, np.complex128)
        cpr = np.zeros((nlev, nt, 9), np.complex128)
        ms = self.member(m)
        mm = lib().orc_edc_chain_proj(C.byref(self.s), C.byref(ms), _p(ep), int(iv), int(nlev), int(project), nt, _p(tn),
                                      _p(tg), _p(twa), _p(a), _p(b), _p(cpr))
        g, _, pv = np_cf_species(a, b, np.zeros(nlev), mm, z, tail, want_pv=True)
        gp = np.einsum("ktl,k->tl", cpr[:mm], pv[:mm])
        m4 = max(1, mm - 4)
        _, _, pv4 = np_cf_species(a, b, np.zeros(nlev), m4, z, tail, want_pv=True)
        gp4 = np.einsum("ktl,k->tl", cpr[:m4], pv4[:m4])
        return g, gp, gp4, (a, b, mm, cpr)

    def lanczos_batch(self, members, iv, nlev):
        nch = len(members)
        ld = int(max(nlev))
        a = np.zeros((nch, ld))
        b = np.zeros((nch, ld))
        mu = np.zeros(nch, np.int32)
        mss = [self.member(m) for m in members]
        arr = (_Member * nch)(*mss)
        ivs = np.ascontiguousarray(iv, dtype=np.int32)
        nl = np.ascontiguousarray(nlev, dtype=np.int32)
        lib().orc_lanczos_batch(C.byref(self.s), arr, _p(ivs), _p(nl), nch, ld, _p(a), _p(b), _p(mu))
        return a, b, mu

    def edc_batch(self, members, eps, iv, nlev, project: bool = True):
        nch = len(members)
        ld = int(max(nlev))
        a = np.zeros((nch, ld), np.complex128)
        b = np.zeros((nch, ld), np.complex128)
        mu = np.zeros(nch, np.int32)
        mss = [self.member(m) for m in members]
        arr = (_Member * nch)(*mss)
        ep = np.ascontiguousarray(eps, dtype=np.complex128)
        ivs = np.ascontiguousarray(iv, dtype=np.int32)
        nl = np.ascontiguousarray(nlev, dtype=np.int32)
        lib().orc_edc_batch(C.byref(self.s), arr, _p(ep), _p(ivs), _p(nl), nch, ld, int(project), _p(a), _p(b), _p(mu))
        return a, b, mu


class OracleReal:
    """Supercell realisation operator (lmto_real_op with per-site species) for the oracle."""

    def __init__(self, nbr, species, c, d, o, e, block, block0):
        self.keep = [np.ascontiguousarray(nbr, dtype=np.int32), np.ascontiguousarray(species, dtype=np.uint8)]
        f = [np.ascontiguousarray(v, dtype=np.float64) for v in (c, d, o, e, block, block0)]
        self.keep += f
        self.dim = 9 * self.keep[0].shape[0]
        self.s = _Real(self.keep[0].shape[0], self.keep[0].shape[1], f[0].shape[0], *[_p(v) for v in self.keep])

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        wk = np.zeros(3 * x.size)
        lib().orc_real_apply(C.byref(self.s), _p(x), _p(y), _p(wk))
        return y

    def kpm_moments(self, c: float, s: float, nmom: int, nvec: int, seed: int, per_vector: bool = False):
        mu = np.zeros(nmom)
        per = np.zeros((nvec, nmom)) if per_vector else None
        lib().orc_kpm_moments(C.byref(self.s), float(c), float(s), int(nmom), int(nvec), int(seed), _p(mu),
                              None if per is None else _p(per))
        return (mu, per) if per_vector else mu

    def lanczos(self, nstep: int, seed: int):
        a = np.zeros(nstep)
        b = np.zeros(nstep)
        lib().orc_real_lanczos(C.byref(self.s), int(nstep), int(seed), _p(a), _p(b))
        return a, b


def threads() -> int:
    """The OpenMP threads of the oracle's parallel loops."""
    return int(lib().orc_threads())


def rademacher(seed: int, v: int, row: int) -> float:
    return lib().orc_rademacher(int(seed), int(v), int(row))


# ---------------------------------------------------------------- numpy restatements of the host-side routines
def np_sqrt_tail(z, ainf, binf):
    w = z - ainf
    s = np.sqrt(w * w - 4.0 * binf * binf + 0j)
    g1 = (w - s) / (2.0 * binf * binf)
    g2 = (w + s) / (2.0 * binf * binf)
    g = np.where(np.abs(g1) <= np.abs(g2), g1, g2)
    return binf * binf * g


def np_cf_green(a, b, n, z, tail):
    g = 1.0 / (z - a[n - 1] - tail)
    for i in range(n - 2, -1, -1):
        g = 1.0 / (z - a[i] - b[i] * b[i] * g)
    return g


def np_bp_asymptote(a, b, n):
    from scipy.linalg import eigvalsh_tridiagonal
    A = 0.5 * (a[n - 1] + a[n - 2])
    B = np.sqrt(0.5 * (b[n - 1] ** 2 + b[n - 2] ** 2))
    for _ in range(500):
        d = np.array(a[:n], dtype=float)
        d[n - 1] = a[n - 1] + B
        ep = eigvalsh_tridiagonal(d, np.array(b[:n - 1], dtype=float), eigvals_only=True)[-1]
        d[n - 1] = a[n - 1] - B
        em = eigvalsh_tridiagonal(d, np.array(b[:n - 1], dtype=float), eigvals_only=True)[0]
        an, bn = 0.5 * (ep + em), 0.25 * (ep - em)
        if abs(an - A) + abs(bn - B) < 1.0e-13:
            return an, bn
        A, B = an, bn
    return A, B


def _chain_matrix(a, b, m, am=None):
    t = np.zeros((m, m), dtype=np.complex128)
    for j in range(m):
        t[j, j] = a[j]
        if j < m - 1:
            t[j, j + 1] = b[j]
            t[j + 1, j] = b[j]
    if am is not None:
        t[m - 1, m - 1] = am
    return t


def np_cf_species(a, b, c, m, z, tail="hull", want_pv=False):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    c = np.asarray(c, dtype=np.complex128)
    info = 1
    if m >= 2 and tail != "none":
        ev = np.linalg.eigvals(_chain_matrix(a, b, m))
        info = 0
    if info == 0:
        ilo = int(np.argmin(ev.real))
        ihi = int(np.argmax(ev.real))
        aa = 0.5 * (ev[ilo] + ev[ihi])
        bb = 0.25 * (ev[ihi] - ev[ilo])
        if tail == "bp":
            for it in range(200):
                ep = np.linalg.eigvals(_chain_matrix(a, b, m, a[m - 1] + bb))
                ep = ep[int(np.argmax(ep.real))]
                em = np.linalg.eigvals(_chain_matrix(a, b, m, a[m - 1] - bb))
                em = em[int(np.argmin(em.real))]
                an, bn = 0.5 * (ep + em), 0.25 * (ep - em)
                if abs(an - aa) + abs(bn - bb) < 1.0e-12 * (1.0 + abs(bn)):
                    aa, bb = an, bn
                    break
                aa, bb = an, bn
    else:
        aa = a[m - 1]
        bb = b[m - 2] if m > 1 else 0.0
    if tail == "avg2" and m >= 3:
        aa = 0.5 * (a[m - 1] + a[m - 2])
        bb = np.sqrt(0.5 * (b[m - 2] ** 2 + b[m - 3] ** 2))
    if tail == "none":
        aa, bb = 0.0, 0.0
    d = z - aa
    if abs(bb) < 1.0e-14:
        tl = 0.0
    else:
        s = np.sqrt(d * d - 4.0 * bb * bb)
        g1 = (d - s) / (2.0 * bb * bb)
        g2 = (d + s) / (2.0 * bb * bb)
        tl = bb * bb * g1 if abs(bb * g1) <= abs(bb * g2) else bb * bb * g2
    gt = np.zeros(m + 1, dtype=np.complex128)
    gt[m - 1] = 1.0 / (z - a[m - 1] - tl)
    for j in range(m - 2, -1, -1):
        gt[j] = 1.0 / (z - a[j] - b[j] ** 2 * gt[j + 1])
    f = 0.0
    prod = gt[0]
    pv = np.zeros(m, np.complex128)
    for k in range(m):
        pv[k] = prod
        f += c[k] * prod
        if k < m - 1:
            prod = prod * b[k] * gt[k + 1]
    if want_pv:
        return gt[0], f, pv
    return gt[0], f


def np_cf_closed(a, b, c, m, z):
    """The fraction of a chain that closed (Krylov space exhausted): exact, no tail; g = G_11, f = sum_k c_k G_k1."""
    gt = np.zeros(m, dtype=np.complex128)
    gt[m - 1] = 1.0 / (z - a[m - 1])
    for j in range(m - 2, -1, -1):
        gt[j] = 1.0 / (z - a[j] - b[j] ** 2 * gt[j + 1])
    f = 0.0
    prod = gt[0]
    for k in range(m):
        f += c[k] * prod
        if k < m - 1:
            prod = prod * b[k] * gt[k + 1]
    return gt[0], f


def np_kpm_dos(mu, c, s, e):
    nmom = len(mu)
    n = np.arange(nmom)
    g = ((nmom - n + 1) * np.cos(np.pi * n / (nmom + 1)) + np.sin(np.pi * n / (nmom + 1)) / np.tan(np.pi / (nmom + 1))) \
        / (nmom + 1)
    e = np.asarray(e, dtype=float)
    x = (e - c) / s
    rho = np.zeros_like(e)
    ok = np.abs(x) < 1.0
    th = np.arccos(x[ok])
    acc = g[0] * mu[0] + 2.0 * np.sum((g[1:] * np.asarray(mu[1:]))[None, :] * np.cos(np.outer(th, n[1:])), axis=1)
    rho[ok] = acc / (np.pi * np.sqrt(1.0 - x[ok] ** 2) * s)
    return rho
