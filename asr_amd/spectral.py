"""Host side of the projected chains for spectral functions (src/asr_spectral.f90, src/asr_edc.f90:166-232).

One chain per centre species, start orbital and energy gives, besides G_c(E) of the start orbital, its projections on the
single-site states of target sites: gp(t, l) = sum_k cpr_k(t, l) pv_k with cpr_k(t, l) = sum_ks tw(t, ks, l) psi_{u_k}(node
ks of t, l) and pv the first column of the terminated chain resolvent, so sum_c x_c gp is the averaged off-site element
<G^(2)_{0 iv, t l}>.  This module holds the dressing weights (spectral_dressing, src/asr_spectral.f90:208-228) and the
unfolding to A(k, E) (spectral_unfold :195-203); the chains run on the GPU (asr_amd.engine.GpuOperator.chains_edc_proj).
"""
from __future__ import annotations

import numpy as np

from . import structure


def spectral_dressing(dc, dl, spc, conc, frozen: bool, nsmax: int) -> np.ndarray:
    """tw (nsmax, 9): tw[k, l] = sum_s x_s T_s(l) phi_k(s), T_s(l) = (dc(l) / dl(s, l))^1/2, for a target site whose class
    has the species spc with concentrations conc (phi: the class basis on the species of non-zero weight, states beyond
    them zero); frozen: the target is the frozen centre (tw = 1 on the reference state).  dc (9,), dl (nsp, 9)."""
    tw = np.zeros((nsmax, 9))
    if frozen:
        tw[0, :] = 1.0
        return tw
    conc = np.asarray(conc, float)
    sup = np.nonzero(conc > 0.0)[0]
    phi = structure.augn_basis(conc[sup])
    dc = np.asarray(dc, float)
    dl = np.asarray(dl, float)
    for k in range(sup.size):
        for j, s in enumerate(sup):
            tw[k] += conc[s] * np.sqrt(dc / dl[spc[s]]) * phi[k, j]
    return tw


def spectral_unfold(disp: np.ndarray, v: np.ndarray, k) -> float:
    """sum_t cos(2 pi k . disp_t) v_t (spectral_unfold, src/asr_spectral.f90:195-203): the Bloch sum of real-space elements
    over target displacements disp (3, nt)."""
    return float(np.sum(np.cos(2.0 * np.pi * (np.asarray(k, float) @ np.asarray(disp, float))) * np.asarray(v, float)))
