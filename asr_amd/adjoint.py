"""Gradients through the real recursion: a torch.autograd harness over the engine's adjoint.

The chain coefficients (a_n, b_n) of real Lanczos chains (lanczos, src/reference/asr_recursion.f90:17-58) are a
differentiable function of the species parameters theta = (c, d) per species and l shell (s, p, d).  `chain_coefficients`
is a torch function of theta whose backward pass is one call of asr_gpu_lanczos_adjoint: the HIP engine recomputes the
chains from Lanczos-vector checkpoints and runs the adjoint three-term recursion, one operator application per level.
Objectives built from (a, b) in torch (the continued fraction on a contour, `contour_objective`) then differentiate end
to end with torch.autograd.

The reference has no adjoint; its nearest analogue is the Hellmann-Feynman trace of the EPI module,
dOmega/dlambda = -(1/pi) Im sum_j w_j Tr[G(z_j) dH/dlambda] (src/asr_epi.f90:14-28), for site-0 parameters only.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from .augcls import Member
from .structure import ORB_L


def members_at(members: list[Member], theta: np.ndarray, c_offset: np.ndarray | None = None) -> list[Member]:
    """Members with c = theta[0] (+ the chain's own offset) and d = theta[1] spread over the orbitals of each l."""
    out = []
    for j, m in enumerate(members):
        c = theta[0] + (c_offset[j] if c_offset is not None else 0.0)
        out.append(dataclasses.replace(m, c=np.ascontiguousarray(c[:, ORB_L]),
                                       d=np.ascontiguousarray(theta[1][:, ORB_L])))
    return out


class _ChainCoefficients(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, gop, members, nlev, c_offset, ckpt):
        th = theta.detach().cpu().numpy().astype(np.float64)
        mem = members_at(members, th, c_offset)
        a, b, mu, _ = gop.lanczos(mem, nlev, tiny=0.0)
        if np.any(mu < nlev):
            raise RuntimeError("chain_coefficients: a chain terminated before its depth")
        ctx.gop, ctx.mem, ctx.nlev, ctx.ckpt = gop, mem, int(nlev), int(ckpt)
        return torch.from_numpy(a).to(theta.device), torch.from_numpy(b).to(theta.device)

    @staticmethod
    def backward(ctx, ga, gb):
        nch = len(ctx.mem)
        la = np.zeros((nch, ctx.nlev)) if ga is None else ga.detach().cpu().numpy()
        lb = np.zeros((nch, ctx.nlev)) if gb is None else gb.detach().cpu().numpy()
        _, _, g, _ = ctx.gop.lanczos_adjoint(ctx.mem, ctx.nlev, la, lb, ckpt=ctx.ckpt)
        return torch.from_numpy(g.sum(axis=0)).to(ga.device if ga is not None else gb.device), None, None, None, None, None


def chain_coefficients(theta: torch.Tensor, gop, members: list[Member], nlev: int, c_offset: np.ndarray | None = None,
                       ckpt: int = 8):
    """(a, b), each nchain x nlev, of the real chains of `members` on the engine operator `gop` at parameters theta
    (2 x nsp x 3 float64: c and d per species and l shell).  Differentiable in theta through the engine's adjoint."""
    return _ChainCoefficients.apply(theta, gop, members, int(nlev), c_offset, int(ckpt))


def cf_green(a: torch.Tensor, b: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
    """Terminated continued fraction G_00(z) of each chain with the avg2 square-root tail (cf_species tail 'avg2',
    src/asr_chain.f90:162-262; the tail of sqrt_tail, src/reference/asr_greens.f90:60-90): a_inf the mean of the last two
    a, b_inf^2 the mean of the last two b^2 that enter the fraction.  a, b: nchain x m (b_{m-1} unused); z: nz complex.
    Returns nchain x nz complex."""
    m = a.shape[1]
    zc = z.to(torch.complex128)[None, :]
    aa = 0.5 * (a[:, m - 1] + a[:, m - 2])
    bb2 = 0.5 * (b[:, m - 2] ** 2 + b[:, m - 3] ** 2)
    d = zc - aa[:, None]
    s = torch.sqrt(d * d - 4.0 * bb2[:, None])
    g1 = (d - s) / (2.0 * bb2[:, None])
    g2 = (d + s) / (2.0 * bb2[:, None])
    # the decaying root (|t| < 1 / b_inf); the branch choice is piecewise constant in the parameters
    pick = (g1.abs() <= g2.abs())
    t = torch.where(pick, g1, g2)
    g = 1.0 / (zc - a[:, m - 1:m] - bb2[:, None] * t)
    for j in range(m - 2, -1, -1):
        g = 1.0 / (zc - a[:, j:j + 1] - (b[:, j:j + 1] ** 2) * g)
    return g


def contour_objective(a: torch.Tensor, b: torch.Tensor, z: torch.Tensor, w: torch.Tensor, weight: torch.Tensor,
                      q0: torch.Tensor, lam: float = 1.0) -> torch.Tensor:
    """Band energy plus a charge penalty on a closed contour (nodes z, weights w: contour_nodes_closed):
    n_c = -(1/pi) Im sum_j w_j G_c(z_j), E_c = -(1/pi) Im sum_j w_j z_j G_c(z_j),
    L = sum_c weight_c E_c + lam/2 sum_c weight_c (n_c - q0_c)^2."""
    g = cf_green(a, b, z)
    wc = w.to(torch.complex128)[None, :]
    n = -(1.0 / np.pi) * torch.imag((wc * g).sum(dim=1))
    e = -(1.0 / np.pi) * torch.imag((wc * z.to(torch.complex128)[None, :] * g).sum(dim=1))
    return (weight * e).sum() + 0.5 * lam * (weight * (n - q0) ** 2).sum()


def chain_dos(a: torch.Tensor, b: torch.Tensor, e, eta: float) -> torch.Tensor:
    """Density of states of each chain at the real energies e with a broadening eta: -(1/pi) Im G_00(e + i eta) of the
    terminated continued fraction (cf_green), nchain x ne; differentiable in a and b, so that d(DOS)/d(theta) comes from
    the engine's adjoint through chain_coefficients."""
    z = torch.as_tensor(np.asarray(e, dtype=np.float64)).to(torch.complex128) + 1j * float(eta)
    return -(1.0 / np.pi) * torch.imag(cf_green(a, b, z))
