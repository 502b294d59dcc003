"""ctypes bindings of libasr_host.so (include/asr_host.h): node-set builder and the host-side chain routines."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
_LIB = None

TAIL = {"hull": 0, "bp": 1, "avg2": 2, "none": 3}


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_PKG, "libasr_host.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with `python -m asr_amd.build host`")
        L = C.CDLL(path)
        P = C.c_void_p
        L.asr_host_build_augcls.argtypes = [C.c_int, P, P, P, C.c_int, P, C.c_int, C.c_int, C.c_int, P, P, C.c_int, C.c_int,
                                            C.POINTER(C.c_void_p), C.c_char_p, C.c_int]
        L.asr_host_tables_dims.argtypes = [C.c_void_p, P]
        L.asr_host_tables_get.argtypes = [C.c_void_p] + [P] * 10
        L.asr_host_tables_free.argtypes = [C.c_void_p]
        L.asr_host_tables_free.restype = None
        L.asr_host_cf_species.argtypes = [P, P, P, C.c_int, P, C.c_int, P, P, P, P]
        L.asr_host_cf_closed.argtypes = [P, P, P, C.c_int, P, P, P]
        L.asr_host_cf_species_batch.argtypes = [C.c_int, P, P, C.c_int, P, P, C.c_int, P]
        L.asr_host_tables_targets.argtypes = [P, C.c_int, P, P, P]
        L.asr_host_bp_asymptote.argtypes = [P, P, C.c_int, P, P, P]
        L.asr_host_terminated_green.argtypes = [P, P, C.c_int, C.c_double, C.c_double, C.c_int, P, P]
        L.asr_host_kpm_dos.argtypes = [P, C.c_int, C.c_double, C.c_double, C.c_int, P, P]
        L.asr_host_tridiag_eigvals.argtypes = [C.c_int, P, P]
        L.asr_host_chain_eigvals.argtypes = [C.c_int, P, P, P]
        L.asr_host_gauss_legendre.argtypes = [C.c_int, P, P]
        L.asr_host_contour_closed.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, P, P]
        L.asr_host_contour_nodes.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double, P, P]
        L.asr_host_fermi_npole.argtypes = [C.c_double]
        L.asr_host_fermi_nodes_count.argtypes = [C.c_int, C.c_int, C.c_int]
        L.asr_host_contour_fermi.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, P, P]
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def build_tables(nbr: np.ndarray, cls: np.ndarray, perm: np.ndarray, start: int, maxlevel: int, nsc, maxrank: int = 0,
                 flippable: np.ndarray | None = None, nthreads: int = 0, tsite=None) -> dict:
    """Node set and node tables (asr_host_build_augcls).  nbr (nsite, nslot) with -1 = no neighbour.  tsite: target sites
    whose single-site nodes are looked up (tnode, tgop: (nt, nsmax), -1 = absent; asr_host_tables_targets)."""
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    nsite, nslot = nbr.shape
    ptr = (np.arange(nsite + 1, dtype=np.int32) * nslot).astype(np.int32)
    flat = np.ascontiguousarray(nbr.reshape(-1))
    cls = np.ascontiguousarray(cls, dtype=np.int32)
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    nsc = np.ascontiguousarray(np.asarray(nsc), dtype=np.int32)
    fl = None if flippable is None else np.ascontiguousarray(flippable, dtype=np.uint8)
    handle = C.c_void_p()
    err = C.create_string_buffer(256)
    st = lib().asr_host_build_augcls(nsite, _p(ptr), _p(flat), _p(cls), perm.shape[0], _p(perm), int(start), int(maxlevel),
                                     int(maxrank), None if fl is None else _p(fl), _p(nsc), nsc.size, int(nthreads),
                                     C.byref(handle), err, 256)
    if st != 0:
        raise RuntimeError(err.value.decode())
    try:
        dims = np.zeros(5, dtype=np.int32)
        lib().asr_host_tables_dims(handle, _p(dims))
        nn, ns, nop, nsm, nstab = (int(v) for v in dims)
        t = dict(hopr=np.zeros((nn, ns), np.int32), hopg=np.zeros((nn, ns), np.int32), flipr=np.zeros((nn, nsm), np.int32),
                 flipg=np.zeros((nn, nsm), np.int32), lst=np.zeros(nn, np.int32), site=np.zeros(nn, np.int32),
                 level=np.zeros(nn, np.int32), orb=np.zeros(nn, np.int32), stab_ptr=np.zeros(nn + 1, np.int32),
                 stab=np.zeros(max(nstab, 1), np.int32))
        lib().asr_host_tables_get(handle, *[_p(t[k]) for k in ("hopr", "hopg", "flipr", "flipg", "lst", "site", "level", "orb",
                                                               "stab_ptr", "stab")])
        t["stab"] = t["stab"][:nstab]
        t["nsmax"] = nsm
        if tsite is not None:
            ts = np.ascontiguousarray(tsite, dtype=np.int32)
            t["tnode"] = np.zeros((ts.size, nsm), np.int32)
            t["tgop"] = np.zeros((ts.size, nsm), np.int32)
            if lib().asr_host_tables_targets(handle, ts.size, _p(ts), _p(t["tnode"]), _p(t["tgop"])) != 0:
                raise RuntimeError("asr_host_tables_targets: target site out of range")
        return t
    finally:
        lib().asr_host_tables_free(handle)


def cf_species(a, b, c, m: int, z: complex, tail: str = "hull", want_pv: bool = False):
    """(g11, f[, pv]) of the terminated complex-symmetric chain (asr_host_cf_species)."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128)[:max(m, 1)])
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128)[:max(m, 1)])
    c = np.ascontiguousarray(np.asarray(c, dtype=np.complex128)[:max(m, 1)])
    zz = np.array([z], dtype=np.complex128)
    g = np.zeros(1, np.complex128)
    f = np.zeros(1, np.complex128)
    pv = np.zeros(max(m, 1), np.complex128)
    lib().asr_host_cf_species(_p(a), _p(b), _p(c), int(m), _p(zz), TAIL[tail], _p(g), _p(f), None, _p(pv))
    if want_pv:
        return g[0], f[0], pv[:m]
    return g[0], f[0]


def cf_species_batch(a, b, m, z, tail: str = "hull") -> np.ndarray:
    """g11 of every chain (rows of a, b; m[j] levels; energies z), cf_species with the same tail, in parallel on the host
    (asr_host_cf_species_batch)."""
    a = np.ascontiguousarray(np.atleast_2d(a), dtype=np.complex128)
    b = np.ascontiguousarray(np.atleast_2d(b), dtype=np.complex128)
    n, ld = a.shape
    m = np.ascontiguousarray(np.broadcast_to(np.asarray(m), (n,)), dtype=np.int32)
    z = np.ascontiguousarray(np.broadcast_to(np.asarray(z, dtype=np.complex128), (n,)))
    g = np.zeros(n, np.complex128)
    lib().asr_host_cf_species_batch(int(n), _p(a), _p(b), int(ld), _p(m), _p(z), TAIL[tail], _p(g))
    return g


def cf_closed(a, b, c, m: int, z: complex):
    """(g11, f) of a complex-symmetric chain that closed: the finite fraction, no tail (asr_host_cf_closed)."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128)[:m])
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128)[:m])
    c = np.ascontiguousarray(np.asarray(c, dtype=np.complex128)[:m])
    zz = np.array([z], dtype=np.complex128)
    g = np.zeros(1, np.complex128)
    f = np.zeros(1, np.complex128)
    if lib().asr_host_cf_closed(_p(a), _p(b), _p(c), int(m), _p(zz), _p(g), _p(f)):
        raise ValueError("cf_closed: m >= 1")
    return g[0], f[0]


def bp_asymptote(a, b, n: int):
    a = np.ascontiguousarray(np.asarray(a, dtype=float))
    b = np.ascontiguousarray(np.asarray(b, dtype=float))
    ai = np.zeros(1)
    bi = np.zeros(1)
    it = np.zeros(1, np.int32)
    st = lib().asr_host_bp_asymptote(_p(a), _p(b), int(n), _p(ai), _p(bi), _p(it))
    if st < 0:
        raise RuntimeError("bp_asymptote: tridiagonal eigenvalue iteration failed")
    return float(ai[0]), float(bi[0]), bool(st == 0)


def terminated_green(a, b, n: int, ainf: float, binf: float, z) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=float))
    b = np.ascontiguousarray(np.asarray(b, dtype=float))
    z = np.ascontiguousarray(np.atleast_1d(np.asarray(z, dtype=np.complex128)))
    g = np.zeros_like(z)
    lib().asr_host_terminated_green(_p(a), _p(b), int(n), float(ainf), float(binf), z.size, _p(z), _p(g))
    return g


def kpm_dos(mu, c: float, s: float, e) -> np.ndarray:
    mu = np.ascontiguousarray(np.asarray(mu, dtype=float))
    e = np.ascontiguousarray(np.asarray(e, dtype=float))
    rho = np.zeros_like(e)
    lib().asr_host_kpm_dos(_p(mu), mu.size, float(c), float(s), e.size, _p(e), _p(rho))
    return rho


def tridiag_eigvals(d, e) -> np.ndarray:
    d = np.ascontiguousarray(np.array(d, dtype=float))
    e = np.ascontiguousarray(np.array(e, dtype=float))
    if lib().asr_host_tridiag_eigvals(d.size, _p(d), _p(e)) != 0:
        raise RuntimeError("tridiag_eigvals did not converge")
    return d


def chain_eigvals(a, b) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128))
    ev = np.zeros(a.size, np.complex128)
    if lib().asr_host_chain_eigvals(a.size, _p(a), _p(b), _p(ev)) != 0:
        raise RuntimeError("chain_eigvals did not converge")
    return ev


def contour_closed(ebot: float, ef: float, eta: float, n_arc: int, n_vert: int):
    z = np.zeros(n_arc + n_vert, np.complex128)
    w = np.zeros(n_arc + n_vert, np.complex128)
    lib().asr_host_contour_closed(float(ebot), float(ef), float(eta), int(n_arc), int(n_vert), _p(z), _p(w))
    return z, w


def gauss_legendre(n: int):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().asr_host_gauss_legendre(int(n), _p(x), _p(w))
    return x, w


def contour_nodes(ebot: float, ef: float, n: int, eta_end: float | None = None):
    z = np.zeros(n, np.complex128)
    w = np.zeros(n, np.complex128)
    if lib().asr_host_contour_nodes(float(ebot), float(ef), int(n), float(eta_end or 0.0), _p(z), _p(w)):
        raise ValueError("contour_nodes: n >= 1 and ef > ebot")
    return z, w


def fermi_npole(kt: float) -> int:
    return int(lib().asr_host_fermi_npole(float(kt)))


def fermi_nodes_count(npole: int, n_main: int, n_step: int) -> int:
    return int(lib().asr_host_fermi_nodes_count(int(npole), int(n_main), int(n_step)))


def contour_fermi(ebot: float, mu: float, kt: float, npole: int, n_main: int, n_step: int):
    n = fermi_nodes_count(npole, n_main, n_step)
    z = np.zeros(n, np.complex128)
    w = np.zeros(n, np.complex128)
    if lib().asr_host_contour_fermi(float(ebot), float(mu), float(kt), int(npole), int(n_main), int(n_step), _p(z), _p(w)):
        raise ValueError("contour_fermi: kt > 0, npole >= 0, n_main >= 1, n_step >= 1")
    return z, w
