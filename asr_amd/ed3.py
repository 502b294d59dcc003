"""Host side of route ed3, the full-ASA energy-dependent recursion (src/asr_ed3.f90).

For a central species c (start site frozen to c) the scaled secular matrix is
    K'(E) = et - h' + u h' + h' u - h' Qt h',   h' = g + D_c^1/2 S D_c^1/2,
    et_s = (D_c / D_s) eps_s,  u_s = o_s eps_s,  Qt_s = (D_s / D_c) (o_s - p_s eps_s),  eps_s = E - E_nu,s,  g_s = (D_c / D_s) c_s,
every coefficient linear in E.  Two complex-symmetric chains on z - K' (from |a> = |u> + o_c h'|u> projected on h'|u>, and
from h'|u> projected on |a>) give the central blocks G_11, (hG)_11, (hGh)_11.

This module holds the species parameters (ed3_setup, src/asr_ed3.f90:53-84), the coefficient arrays in the form the chain
engine takes (ed3_coefs :87-97 at one energy, the linear form cf(l, s, 1:6) of ed3_blocks_blk :296-303) and the per-chain
operator parameters.  The chains themselves run on the GPU (asr_amd.engine.GpuOperator.ed3_blocks).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .augcls import Member

ORB_L = np.array([0, 1, 1, 1, 2, 2, 2, 2, 2])


def per_orbital(v_l: np.ndarray) -> np.ndarray:
    """(nsp, 3) per l -> (nsp, 9) per orbital (s, p x 3, d x 5)."""
    return np.asarray(v_l, dtype=float)[:, ORB_L]


@dataclass
class Ed3Species:
    """ed3_t: D_c (9); D_s, E_nu,s, o_s, p_s, g_s (nsp, 9) for the central species `central`."""
    central: int
    dc: np.ndarray
    dl: np.ndarray
    enu: np.ndarray
    o: np.ndarray
    p: np.ndarray
    g: np.ndarray
    sqrdel_c: np.ndarray


def ed3_setup(pp: dict, central: int) -> Ed3Species:
    """Species parameters of K' for the central species (ed3_setup, src/asr_ed3.f90:53-84).  pp holds per species and l
    (nsp, 3), in the screened (alpha) representation: c = C - E_nu, sqrdel (sqrt Delta), enu, o, p."""
    sd = per_orbital(pp["sqrdel"])
    dl = sd ** 2
    dc = dl[central].copy()
    g = dc[None, :] / dl * per_orbital(pp["c"])
    return Ed3Species(central=int(central), dc=dc, dl=dl, enu=per_orbital(pp["enu"]), o=per_orbital(pp["o"]),
                      p=per_orbital(pp["p"]), g=g, sqrdel_c=sd[central].copy())


def ed3_coefs(ed: Ed3Species, z: complex):
    """et, u, Qt (nsp, 9) at z (ed3_coefs, src/asr_ed3.f90:87-97)."""
    eps = z - ed.enu
    et = ed.dc[None, :] / ed.dl * eps
    uu = ed.o * eps
    qt = ed.dl / ed.dc[None, :] * (ed.o - ed.p * eps)
    return et, uu, qt


def ed3_linear(ed: Ed3Species) -> np.ndarray:
    """cf (6, nsp, 9) with et = z cf1 + cf2, u = z cf3 + cf4, Qt = z cf5 + cf6 (src/asr_ed3.f90:296-303)."""
    r = ed.dc[None, :] / ed.dl
    cf = np.empty((6,) + ed.dl.shape)
    cf[0] = r
    cf[1] = -r * ed.enu
    cf[2] = ed.o
    cf[3] = -ed.o * ed.enu
    cf[4] = -ed.dl / ed.dc[None, :] * ed.p
    cf[5] = ed.dl / ed.dc[None, :] * (ed.o + ed.p * ed.enu)
    return cf


def ed3_member(ed: Ed3Species, l0: int, z: complex) -> Member:
    """Operator parameters of one ed3 chain: h' = c + d S d with c = g, d = D_c^1/2 for every species, o = e = 0, start site
    frozen to the central species (the parameters ed3_setup puts into the class operator)."""
    nsp = ed.g.shape[0]
    d = np.broadcast_to(np.abs(ed.sqrdel_c), (nsp, 9)).copy()
    zero = np.zeros((nsp, 9))
    return Member(l0=int(l0), c=ed.g.copy(), d=d, o=zero, e=zero.copy(), frozen=ed.central, z=complex(z))


def ed3_oc(ed: Ed3Species, l0: int) -> float:
    """o_c of the start orbital (the mixing of |a> = |u> + o_c |v>)."""
    return float(ed.o[ed.central, l0])


def ed3_params(rng: np.random.Generator, nsp: int, nspin: int):
    """Seeded species parameters per spin in the alpha representation: c = C - E_nu ~ U[-0.5, 0.5], sqrt(Delta) ~
    U[0.05, 0.15], E_nu ~ U[-0.1, 0.1], o ~ U[-0.05, 0.05], p = o^2 + p^gamma with p^gamma ~ U[0, 0.02] (Ry units)."""
    out = []
    for _ in range(nspin):
        c = rng.uniform(-0.5, 0.5, (nsp, 3))
        sd = rng.uniform(0.05, 0.15, (nsp, 3))
        enu = rng.uniform(-0.1, 0.1, (nsp, 3))
        o = rng.uniform(-0.05, 0.05, (nsp, 3))
        p = o ** 2 + rng.uniform(0.0, 0.02, (nsp, 3))
        out.append(dict(c=c, sqrdel=sd, enu=enu, o=o, p=p))
    return out
