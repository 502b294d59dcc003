"""Python side of the MI355X chain engine: ctypes bindings of libasr_hip.so (include/asr_hip.h).

This mirrors the reference's operator interface for the hot path (augcls_op with augcls_set_params / set_start /
freeze_site / set_class_conc, src/asr_augcls.f90:107-185; the chain routines lanczos, edc_green, edc_green_blk; kpm_moments):
the operator tables live on the GPU, a chain is described by its start orbital, species parameters, frozen species and
(for route edc) its energy.  There is no CPU fallback: without the HIP library every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .augcls import AugclsOperator, Member

_PKG = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class EngineError(RuntimeError):
    pass


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_PKG, "libasr_hip.so")
        if not os.path.exists(path):
            raise EngineError(f"{path} is missing: build it with `python -m asr_amd.build hip`")
        L = C.CDLL(path)
        P = C.c_void_p
        L.asr_gpu_last_error.restype = C.c_char_p
        L.asr_gpu_device_count.argtypes = [P]
        L.asr_gpu_op_create.argtypes = [C.c_int] * 7 + [C.POINTER(C.c_void_p)]
        L.asr_gpu_op_set_tables.argtypes = [P] * 11 + [P, C.c_int]
        L.asr_gpu_op_set_orbital_ops.argtypes = [P, P, P]
        L.asr_gpu_op_set_blocks.argtypes = [P, P, P]
        L.asr_gpu_op_set_classes.argtypes = [P, P, P, P]
        L.asr_gpu_op_set_tile.argtypes = [P, C.c_int]
        L.asr_gpu_op_set_depth_tol.argtypes = [P, C.c_double, C.c_int]
        L.asr_gpu_op_set_reorth.argtypes = [P, C.c_int]
        L.asr_gpu_op_set_level_form.argtypes = [P, C.c_int]
        L.asr_gpu_op_set_compaction.argtypes = [P, C.c_int]
        L.asr_gpu_op_get_compactions.argtypes = [P, C.POINTER(C.c_int)]
        L.asr_gpu_op_set_window.argtypes = [P, C.c_int]
        L.asr_gpu_op_set_window_queue.argtypes = [P, C.c_int]
        L.asr_gpu_op_set_targets.argtypes = [P, C.c_int, P, P]
        L.asr_gpu_chains_edc_proj.argtypes = [P, C.c_int, P, P, P, P, P, P, P, C.c_int, C.c_int, P, P, P, P, P, P, P, C.c_int, P]
        L.asr_gpu_op_get_tile.argtypes = [P, C.c_int, C.c_int, P]
        L.asr_gpu_real_set_tile.argtypes = [P, C.c_int]
        L.asr_gpu_real_set_doubling.argtypes = [P, C.c_int]
        L.asr_gpu_op_trim.argtypes = [P]
        L.asr_gpu_op_destroy.argtypes = [P]
        L.asr_gpu_op_destroy.restype = None
        L.asr_gpu_apply_batched.argtypes = [P, C.c_int, P, P, P, P, P, P]
        L.asr_gpu_apply_complex.argtypes = [P, C.c_int, P, P, P, P, P, P, P, P]
        L.asr_gpu_lanczos_real.argtypes = [P, C.c_int, P, P, P, P, P, C.c_double, C.c_int, P, P, P, C.c_int, P]
        L.asr_gpu_chains_edc.argtypes = [P, C.c_int, P, P, P, P, P, P, P, C.c_int, C.c_int, P, P, P, C.c_int, P]
        L.asr_gpu_chains_edc_multi.argtypes = [C.c_int, P, P, P, P, P, P, P, P, P, C.c_int, C.c_int, P, P, P, C.c_int, P]
        L.asr_gpu_lanczos_real_multi.argtypes = [C.c_int, P, P, P, P, P, P, P, C.c_double, C.c_int, P, P, P, C.c_int, P]
        L.asr_gpu_ed3_blocks.argtypes = [P, C.c_int, P, P, P, P, P, P, P, P, C.c_int, C.c_int, P, P, P, P, P, C.c_int, P]
        L.asr_gpu_epi_blocks.argtypes = [P, C.c_int, P, P, P, P, P, P, P, P, P, C.c_int, C.c_int, P, P, P, P, P, C.c_int, P]
        L.asr_gpu_lanczos_adjoint.argtypes = [P, C.c_int, P, P, P, P, C.c_int, C.c_int, P, P, P, P, P, P, P]
        L.asr_gpu_profile_chains.argtypes = [P, C.c_int, P, P, P, P, P, P, P, C.c_int, P, P, P]
        L.asr_gpu_real_create.argtypes = [C.c_int, C.c_int, P, C.c_int, P, P, P, P, C.POINTER(C.c_void_p)]
        L.asr_gpu_real_destroy.argtypes = [P]
        L.asr_gpu_real_destroy.restype = None
        L.asr_gpu_real_apply.argtypes = [P, C.c_int, P, P]
        L.asr_gpu_spectral_bounds.argtypes = [P, C.c_int, C.c_int, C.c_uint64, C.c_double, P, P, P, P]
        L.asr_gpu_kpm_moments.argtypes = [P, C.c_double, C.c_double, C.c_int, C.c_int, C.c_uint64, P, P, P, P]
        L.asr_gpu_kpm_moments_noalloc.argtypes = [P, C.c_double, C.c_double, C.c_int, C.c_int, C.c_uint64, P]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(status: int, what: str):
    if status != 0:
        raise EngineError(f"{what}: {lib().asr_gpu_last_error().decode()}")


def device_count() -> int:
    n = C.c_int(0)
    _check(lib().asr_gpu_device_count(C.byref(n)), "asr_gpu_device_count")
    return n.value


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def pack_members(op: AugclsOperator, members: list[Member], complex_route: bool = False):
    nb = len(members)
    start = _i32([m.l0 for m in members])
    params = _f64(np.stack([np.stack([np.asarray(m.c, float).reshape(op.nsp, 9), np.asarray(m.d, float).reshape(op.nsp, 9),
                                      np.asarray(m.o, float).reshape(op.nsp, 9), np.asarray(m.e, float).reshape(op.nsp, 9)])
                            for m in members]))
    frozen = _i32([m.frozen for m in members])
    conc = None
    if any(m.conc is not None for m in members):
        conc = np.zeros((nb, op.ncls, op.nsmax))
        for j, m in enumerate(members):
            cc = op.conc if m.conc is None else m.conc
            for kap in range(op.ncls):
                conc[j, kap, :op.nsc[kap]] = np.asarray(cc[kap])[:op.nsc[kap]]
        conc = _f64(conc)
    pab = z = None
    if complex_route:
        pab = _f64(np.stack([np.stack([np.asarray(m.pa, float).reshape(op.nsp, 9), np.asarray(m.pb, float).reshape(op.nsp, 9)])
                             for m in members]))
        z = np.ascontiguousarray(np.array([complex(m.z) for m in members], dtype=np.complex128))
    return start, params, frozen, conc, pab, z


class GpuOperator:
    """An augcls operator resident on the GPU (one node set, shared by every chain on it)."""

    def __init__(self, op: AugclsOperator):
        self.op = op
        h = C.c_void_p()
        _check(lib().asr_gpu_op_create(op.nn, op.nslot, op.nop, op.nsmax, op.ncls, op.nsp, op.ntype, C.byref(h)),
               "asr_gpu_op_create")
        self.h = h
        stab = op.stab if op.stab.size else np.zeros(1, np.int32)
        _check(lib().asr_gpu_op_set_tables(h, _p(_i32(op.hopr)), _p(_i32(op.hopg)), _p(_i32(op.flipr)), _p(_i32(op.flipg)),
                                           _p(_i32(op.lst)), _p(_i32(op.ncl)), _p(_i32(op.ntyp)),
                                           _p(np.ascontiguousarray(op.fz, np.uint8)), _p(_f64(op.sq)), _p(_i32(op.stab_ptr)),
                                           _p(_i32(stab)), int(op.stab.size)), "asr_gpu_op_set_tables")
        _check(lib().asr_gpu_op_set_orbital_ops(h, _p(_i32(op.pi)), _p(_f64(op.sg))), "asr_gpu_op_set_orbital_ops")
        _check(lib().asr_gpu_op_set_blocks(h, _p(_f64(op.block)), _p(_f64(op.block0))), "asr_gpu_op_set_blocks")
        _check(lib().asr_gpu_op_set_classes(h, _p(_i32(op.nsc)), _p(_i32(op.spc)), _p(_f64(op.mt))), "asr_gpu_op_set_classes")

    def close(self):
        if self.h:
            lib().asr_gpu_op_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tile(self, mt: int):
        """Member tile of the batched chains (power of two <= 64; 0 = autotuned)."""
        _check(lib().asr_gpu_op_set_tile(self.h, int(mt)), "asr_gpu_op_set_tile")

    def set_targets(self, tnode: np.ndarray, tgop: np.ndarray):
        """Target states of the projected chains (tnode / tgop: (ntarget, nsmax), -1 = none)."""
        tn, tg = _i32(tnode), _i32(tgop)
        _check(lib().asr_gpu_op_set_targets(self.h, tn.shape[0], _p(tn), _p(tg)), "asr_gpu_op_set_targets")
        self.ntarget = tn.shape[0]

    def chains_edc_proj(self, members: list[Member], nlev, tw, project: bool = True, mode: int = 0):
        """Projected chains (edc_green_proj) on the target states of set_targets: tw (nchain, ntarget, nsmax, 9)
        dressing weights.  Returns g (nchain,), gp, gp4 (nchain, ntarget, 9), a, b, mused and the device time in ms."""
        nch = len(members)
        nlev = _i32(np.broadcast_to(np.asarray(nlev), (nch,)))
        ld = int(nlev.max())
        tw = _f64(tw)
        nt = tw.shape[1]
        a = np.zeros((nch, ld), np.complex128)
        b = np.zeros((nch, ld), np.complex128)
        mu = np.zeros(nch, np.int32)
        g = np.zeros(nch, np.complex128)
        gp = np.zeros((nch, nt, 9), np.complex128)
        gp4 = np.zeros((nch, nt, 9), np.complex128)
        ms = np.zeros(1)
        s, p, f, c, pab, z = pack_members(self.op, members, complex_route=True)
        _check(lib().asr_gpu_chains_edc_proj(self.h, nch, _p(s), _p(p), _p(pab), _p(f), _p(c), _p(z), _p(nlev), int(project),
                                             ld, _p(tw), _p(a), _p(b), _p(mu), _p(g), _p(gp), _p(gp4), int(mode), _p(ms)),
               "asr_gpu_chains_edc_proj")
        return g, gp, gp4, a, b, mu, float(ms[0])

    def trim(self):
        """Release the level-vector blocks cached between batched runs."""
        _check(lib().asr_gpu_op_trim(self.h), "asr_gpu_op_trim")

    def set_reorth(self, reorth: bool):
        """Full reorthogonalisation of the real chains (lanczos reorth = .true.; small cases)."""
        _check(lib().asr_gpu_op_set_reorth(self.h, int(bool(reorth))), "asr_gpu_op_set_reorth")

    def set_level_form(self, form: int):
        """Form of an edc chain level (and of the chain blocks of route ed3 and EPI): 0 the fused level (a_n from the first
        operator pass, the update folded into the second), 1 the separate update pass in the reference's order
        (edc_green, chain).  For the adjoint: 0 the recomputed level and u_bar on the operator's last pass, 1 in passes of
        their own."""
        _check(lib().asr_gpu_op_set_level_form(self.h, int(form)), "asr_gpu_op_set_level_form")

    def set_compaction(self, on: bool):
        """Member compaction of batched chains with the adaptive depth (host-issued levels): the running members move to
        the front after each level of the depth test, so that member tiles whose members have left stop."""
        _check(lib().asr_gpu_op_set_compaction(self.h, int(bool(on))), "asr_gpu_op_set_compaction")

    def set_window(self, nslot: int):
        """Window runs of the batched edc chains with the adaptive depth on the device: nslot member slots, the other
        chains queued for the slots of members that leave (0: off)."""
        _check(lib().asr_gpu_op_set_window(self.h, int(nslot)), "asr_gpu_op_set_window")

    def set_window_queue(self, on_device: bool):
        """Where a window run's queue lives: on the device (the default; a kernel after each level), or followed by the host
        after every level."""
        _check(lib().asr_gpu_op_set_window_queue(self.h, int(bool(on_device))), "asr_gpu_op_set_window_queue")

    def compactions(self) -> int:
        """Member compactions of the last batched run."""
        n = C.c_int(0)
        _check(lib().asr_gpu_op_get_compactions(self.h, C.byref(n)), "asr_gpu_op_get_compactions")
        return int(n.value)

    def set_depth_tol(self, depth_tol: float, tail: str = "hull"):
        """Adaptive depth of the edc chains (ed_depth_tol / ed_tail of edc_green); 0 = every chain to its depth."""
        tails = {"hull": 0, "bp": 1, "avg2": 2, "none": 3}
        _check(lib().asr_gpu_op_set_depth_tol(self.h, float(depth_tol), tails[tail]), "asr_gpu_op_set_depth_tol")

    def tuned_tile(self, nchain: int, is_complex: bool) -> int:
        mt = np.zeros(1, np.int32)
        _check(lib().asr_gpu_op_get_tile(self.h, int(nchain), int(is_complex), _p(mt)), "asr_gpu_op_get_tile")
        return int(mt[0])

    def apply(self, members: list[Member], x: np.ndarray) -> np.ndarray:
        """y = H x for every member (x: nb x 9 nn, real)."""
        x = _f64(np.atleast_2d(x))
        if x.shape != (len(members), self.op.dim):
            raise ValueError("apply: x must be (members, 9 nn)")
        y = np.zeros_like(x)
        s, p, f, c, _, _ = pack_members(self.op, members)
        _check(lib().asr_gpu_apply_batched(self.h, len(members), _p(s), _p(p), _p(f), _p(c), _p(x), _p(y)),
               "asr_gpu_apply_batched")
        return y

    def apply_complex(self, members: list[Member], x: np.ndarray) -> np.ndarray:
        """y = H x + X(z) x for complex members (route edc operator)."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.complex128)
        y = np.zeros_like(x)
        s, p, f, c, pab, z = pack_members(self.op, members, complex_route=True)
        _check(lib().asr_gpu_apply_complex(self.h, len(members), _p(s), _p(p), _p(pab), _p(f), _p(c), _p(z), _p(x), _p(y)),
               "asr_gpu_apply_complex")
        return y

    def lanczos(self, members: list[Member], nlev, tiny: float = 1.0e-12, mode: int = 0):
        """Real Lanczos chains (lanczos) of every member; mode 0 batched, 1 plain (one chain per launch)."""
        nch = len(members)
        nlev = _i32(np.broadcast_to(np.asarray(nlev), (nch,)))
        ld = int(nlev.max())
        a = np.zeros((nch, ld))
        b = np.zeros((nch, ld))
        mu = np.zeros(nch, np.int32)
        ms = np.zeros(1)
        s, p, f, c, _, _ = pack_members(self.op, members)
        _check(lib().asr_gpu_lanczos_real(self.h, nch, _p(s), _p(p), _p(f), _p(c), _p(nlev), float(tiny), ld, _p(a), _p(b),
                                          _p(mu), int(mode), _p(ms)), "asr_gpu_lanczos_real")
        return a, b, mu, float(ms[0])

    def lanczos_adjoint(self, members: list[Member], nlev: int, dl_da: np.ndarray, dl_db: np.ndarray, ckpt: int = 8):
        """Forward real chains of depth nlev and the gradient of L(a, b) with respect to (c, d) of every species and l shell.

        dl_da, dl_db: nchain x nlev.  Returns a, b (nchain x nlev), grad (nchain x 2 x nsp x 3: dL/dc, dL/dd) and the device
        times (forward ms, reverse ms)."""
        nch = len(members)
        nlev = int(nlev)
        la = _f64(np.broadcast_to(np.asarray(dl_da, np.float64), (nch, nlev)))
        lb = _f64(np.broadcast_to(np.asarray(dl_db, np.float64), (nch, nlev)))
        a = np.zeros((nch, nlev))
        b = np.zeros((nch, nlev))
        g = np.zeros((nch, 2, self.op.nsp, 3))
        fm = np.zeros(1)
        am = np.zeros(1)
        s, p, f, c, _, _ = pack_members(self.op, members)
        _check(lib().asr_gpu_lanczos_adjoint(self.h, nch, _p(s), _p(p), _p(f), _p(c), nlev, int(ckpt), _p(la), _p(lb), _p(a),
                                             _p(b), _p(g), _p(fm), _p(am)), "asr_gpu_lanczos_adjoint")
        return a, b, g, (float(fm[0]), float(am[0]))

    def chains_edc(self, members: list[Member], nlev, project: bool = True, mode: int = 0):
        """Complex-symmetric chains of route edc (edc_green / edc_green_blk, depth_tol = 0)."""
        nch = len(members)
        nlev = _i32(np.broadcast_to(np.asarray(nlev), (nch,)))
        ld = int(nlev.max())
        a = np.zeros((nch, ld), np.complex128)
        b = np.zeros((nch, ld), np.complex128)
        mu = np.zeros(nch, np.int32)
        ms = np.zeros(1)
        s, p, f, c, pab, z = pack_members(self.op, members, complex_route=True)
        _check(lib().asr_gpu_chains_edc(self.h, nch, _p(s), _p(p), _p(pab), _p(f), _p(c), _p(z), _p(nlev), int(project), ld,
                                        _p(a), _p(b), _p(mu), int(mode), _p(ms)), "asr_gpu_chains_edc")
        return a, b, mu, float(ms[0])

    def ed3_blocks(self, members: list[Member], cf, oc, nlev, project: bool = True, mode: int = 0):
        """Route ed3 (ed3_blocks / ed3_blocks_blk): members from asr_amd.ed3.ed3_member (c = g, d = D_c^1/2, o = e = 0,
        frozen central species, z), cf (nchain, 6, nsp, 9) from ed3_linear, oc (nchain) = o_c of the start orbital.
        Returns bk (nchain, 4) = (G_11, (hG)_11, (hGh)_11, q1a - q1b), the chains' a, b, c (nchain, 2, ld), mused
        (nchain, 2) and the device time in ms.  mode 0 batched, 1 plain, 2 one HIP graph."""
        nch = len(members)
        nlev = _i32(np.broadcast_to(np.asarray(nlev), (nch,)))
        ld = int(nlev.max())
        cf = _f64(np.broadcast_to(np.asarray(cf, np.float64), (nch, 6, self.op.nsp, 9)))
        oc = _f64(np.broadcast_to(np.asarray(oc, np.float64), (nch,)))
        bk = np.zeros((nch, 4), np.complex128)
        a = np.zeros((nch, 2, ld), np.complex128)
        b = np.zeros((nch, 2, ld), np.complex128)
        c = np.zeros((nch, 2, ld), np.complex128)
        mu = np.zeros((nch, 2), np.int32)
        ms = np.zeros(1)
        s, p, f, cc, _, _ = pack_members(self.op, members)
        z = np.ascontiguousarray(np.array([complex(m.z) for m in members], dtype=np.complex128))
        _check(lib().asr_gpu_ed3_blocks(self.h, nch, _p(s), _p(p), _p(cf), _p(oc), _p(f), _p(cc), _p(z), _p(nlev),
                                        int(project), ld, _p(bk), _p(a), _p(b), _p(c), _p(mu), int(mode), _p(ms)),
               "asr_gpu_ed3_blocks")
        return bk, a, b, c, mu, float(ms[0])

    def epi_blocks(self, members: list[Member], v9, nlev, project: bool = True, mode: int = 0):
        """EPI blocks (epi_blocks / epi_blocks_blk): members from asr_amd.epi.epi_member (edc parameters of epi_setup, site
        0 frozen to the path point, frozen2 the partner's species, chi the characters of the start vector, z), v9 (nchain, 9)
        unit start combinations on the start node.  Returns x (nchain, 4) = (X0, X1, X2, X1a - X1b), the chains' a, b, c
        (nchain, 2, ld), mused (nchain, 2) and the device time in ms."""
        nch = len(members)
        nlev = _i32(np.broadcast_to(np.asarray(nlev), (nch,)))
        ld = int(nlev.max())
        v9 = _f64(np.broadcast_to(np.asarray(v9, np.float64), (nch, 9)))
        chi = _f64(np.stack([np.asarray(m.chi, np.float64).reshape(self.op.nop) for m in members]))
        fz2 = _i32([m.frozen2 for m in members])
        x = np.zeros((nch, 4), np.complex128)
        a = np.zeros((nch, 2, ld), np.complex128)
        b = np.zeros((nch, 2, ld), np.complex128)
        c = np.zeros((nch, 2, ld), np.complex128)
        mu = np.zeros((nch, 2), np.int32)
        ms = np.zeros(1)
        _, p, f, cc, pab, z = pack_members(self.op, members, complex_route=True)
        _check(lib().asr_gpu_epi_blocks(self.h, nch, _p(v9), _p(chi), _p(p), _p(pab), _p(f), _p(fz2), _p(cc), _p(z),
                                        _p(nlev), int(project), ld, _p(x), _p(a), _p(b), _p(c), _p(mu), int(mode), _p(ms)),
               "asr_gpu_epi_blocks")
        return x, a, b, c, mu, float(ms[0])

    def profile(self, members: list[Member], nlev, project: bool = True, complex_route: bool = True):
        """Mean duration (ms) and kind of every kernel launch of one chain level."""
        nch = len(members)
        nlev = _i32(np.broadcast_to(np.asarray(nlev), (nch,)))
        s, p, f, c, pab, z = pack_members(self.op, members, complex_route=complex_route)
        nk = np.zeros(1, np.int32)
        ms = np.zeros(16)
        kind = np.zeros(16, np.int32)
        _check(lib().asr_gpu_profile_chains(self.h, nch, _p(s), _p(p), _p(pab), _p(f), _p(c), _p(z), _p(nlev), int(project),
                                            _p(nk), _p(ms), _p(kind)), "asr_gpu_profile_chains")
        n = int(nk[0])
        return ms[:n].copy(), kind[:n].copy()


def chains_multi(groups, project: bool = True, mode: int = 2, complex_route: bool = True, tiny: float = 1.0e-12):
    """Chains on several operators in one batched run.  groups: list of (GpuOperator, members, nlev).  mode 2: one HIP
    graph for the whole run, 0: host-issued levels on one stream per group, 1: plain (one chain per launch sequence).
    Returns per group (a, b, mused) and the device time in ms."""
    ng = len(groups)
    ops = (C.c_void_p * ng)(*[g[0].h.value for g in groups])
    nch = _i32([len(g[1]) for g in groups])
    packs = [pack_members(g[0].op, g[1], complex_route=complex_route) for g in groups]
    start = _i32(np.concatenate([p[0] for p in packs]))
    params = _f64(np.concatenate([p[1].reshape(-1) for p in packs]))
    frozen = _i32(np.concatenate([p[2] for p in packs]))
    conc = None
    if any(p[3] is not None for p in packs):
        parts = []
        for (G, mem, _), p in zip(groups, packs):
            if p[3] is not None:
                parts.append(p[3].reshape(-1))
            else:
                c = np.zeros((len(mem), G.op.ncls, G.op.nsmax))
                for kap in range(G.op.ncls):
                    c[:, kap, :G.op.nsc[kap]] = G.op.conc[kap, :G.op.nsc[kap]]
                parts.append(c.reshape(-1))
        conc = _f64(np.concatenate(parts))
    nlev = _i32(np.concatenate([np.broadcast_to(np.asarray(g[2]), (len(g[1]),)) for g in groups]))
    ld = int(nlev.max())
    tot = int(nch.sum())
    mu = np.zeros(tot, np.int32)
    ms = np.zeros(1)
    if complex_route:
        pab = _f64(np.concatenate([p[4].reshape(-1) for p in packs]))
        z = np.ascontiguousarray(np.concatenate([p[5] for p in packs]))
        a = np.zeros((tot, ld), np.complex128)
        b = np.zeros((tot, ld), np.complex128)
        _check(lib().asr_gpu_chains_edc_multi(ng, ops, _p(nch), _p(start), _p(params), _p(pab), _p(frozen), _p(conc), _p(z),
                                              _p(nlev), int(project), ld, _p(a), _p(b), _p(mu), int(mode), _p(ms)),
               "asr_gpu_chains_edc_multi")
    else:
        a = np.zeros((tot, ld))
        b = np.zeros((tot, ld))
        _check(lib().asr_gpu_lanczos_real_multi(ng, ops, _p(nch), _p(start), _p(params), _p(frozen), _p(conc), _p(nlev),
                                                float(tiny), ld, _p(a), _p(b), _p(mu), int(mode), _p(ms)),
               "asr_gpu_lanczos_real_multi")
    out = []
    off = 0
    for n in nch:
        out.append((a[off:off + n], b[off:off + n], mu[off:off + n]))
        off += int(n)
    return out, float(ms[0])


class GpuRealization:
    """Supercell realisation h = c + d S d on the GPU (KPM operator)."""

    def __init__(self, nbr: np.ndarray, species: np.ndarray, par: np.ndarray, block: np.ndarray, block0: np.ndarray):
        nbr = _i32(nbr)
        self.nsite, self.nslot = nbr.shape
        self.dim = 9 * self.nsite
        par = _f64(par)
        nsp = par.shape[1]
        h = C.c_void_p()
        _check(lib().asr_gpu_real_create(self.nsite, self.nslot, _p(nbr), nsp, _p(np.ascontiguousarray(species, np.uint8)),
                                         _p(par), _p(_f64(block)), _p(_f64(block0)), C.byref(h)), "asr_gpu_real_create")
        self.h = h

    def close(self):
        if self.h:
            lib().asr_gpu_real_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tile(self, vt: int):
        _check(lib().asr_gpu_real_set_tile(self.h, int(vt)), "asr_gpu_real_set_tile")

    def set_doubling(self, on: bool):
        """kpm_moments by moment doubling (two moments per Chebyshev step; equal to the direct moments up to rounding)."""
        _check(lib().asr_gpu_real_set_doubling(self.h, int(bool(on))), "asr_gpu_real_set_doubling")

    def apply(self, x: np.ndarray) -> np.ndarray:
        x = _f64(np.atleast_2d(x))
        y = np.zeros_like(x)
        _check(lib().asr_gpu_real_apply(self.h, x.shape[0], _p(x), _p(y)), "asr_gpu_real_apply")
        return y

    def spectral_bounds(self, nstep: int = 80, margin: float = 0.0, nvec: int = 1, seed: int = 7):
        """(emin, emax, a, b): extremal eigenvalues of h from nstep Lanczos steps of nvec random vectors, widened by
        margin (spectral_bounds); a, b: (nvec, nstep) chain coefficients."""
        lo, hi = np.zeros(1), np.zeros(1)
        a = np.zeros((nvec, nstep))
        b = np.zeros((nvec, nstep))
        _check(lib().asr_gpu_spectral_bounds(self.h, int(nstep), int(nvec), int(seed), float(margin), _p(lo), _p(hi), _p(a),
                                             _p(b)), "asr_gpu_spectral_bounds")
        return float(lo[0]), float(hi[0]), a, b

    def kpm_moments_noalloc(self, c: float, s: float, nmom: int, nvec: int, seed: int, mu_dev: int):
        """The moments into nmom doubles of device memory at address mu_dev (asr_gpu_kpm_moments_noalloc: the work arrays
        in the library's fixed device scratch, no device allocation)."""
        _check(lib().asr_gpu_kpm_moments_noalloc(self.h, float(c), float(s), int(nmom), int(nvec), int(seed),
                                                 C.c_void_p(int(mu_dev))), "asr_gpu_kpm_moments_noalloc")

    def kpm_moments(self, c: float, s: float, nmom: int, nvec: int, seed: int, per_vector: bool = False):
        mu = np.zeros(nmom)
        per = np.zeros((nvec, nmom)) if per_vector else None
        step = np.zeros(1)
        tot = np.zeros(1)
        _check(lib().asr_gpu_kpm_moments(self.h, float(c), float(s), int(nmom), int(nvec), int(seed), _p(mu), _p(per),
                                         _p(step), _p(tot)), "asr_gpu_kpm_moments")
        out = dict(mu=mu, step_ms=float(step[0]), total_ms=float(tot[0]))
        if per_vector:
            out["mu_per"] = per
        return out
