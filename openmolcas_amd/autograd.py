"""Differentiable closed-shell energy of the Cholesky J/K build.

    E(C, L, h) = tr(P h) + 1/2 sum_J V_J^2 - 1/4 sum_J tr(P L^J P L^J)
    P = 2 C C^T,  V_J = tr(P L^J)
             = tr(P h) + 1/2 tr(P J[P]) - 1/4 tr(P K[P])

The forward pass is the J/K build (Coulomb streaming + exchange half-transform
and SYRK) followed by the FP64 energy step; the backward pass is hand written
in the HIP library (mjk_grad_orbitals, mjk_grad_vectors):
    dE/dC = 4 F C,  F = h + J[P] - 1/2 K[P]
    dE/dL^J (packed) = s_ab (V_J P_ab - 1/2 (P L^J P)_ab),  s = 2 off the diagonal
    dE/dh (packed)   = fold(P)
L is the (nchol, npair) tensor of packed vectors; it is uploaded to the
context's HBM store when it differs from the resident copy.
"""
from __future__ import annotations

import torch

from .fock import JKContext


class ClosedShellEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, C, L, h, jk: JKContext, precision="fp64"):
        jk.ensure_vectors(L)
        C = C.detach().to(torch.float64).contiguous()
        h = h.detach().to(torch.float64).reshape(1, -1).contiguous()
        dlt = jk.density(C, scale=2.0)              # fold(P)
        if precision == "fp64":                     # V pass inside the half-transform
            J, K, V = jk.coulomb_exchange(dlt, [C], want_v=True)
        else:
            J, V = jk.coulomb(dlt, want_v=True)
            K = jk.exchange([C], precision=precision)   # K[C C^T] = 1/2 K[P]
        G = torch.zeros_like(J)
        jk.fock_update(G, J, [0], [1.0], K, [-1.0])  # G = J - 1/2 K[P]
        E = jk.energy(h, dlt, G)
        F = h + G
        ctx.jk = jk
        ctx.save_for_backward(C, F, V[:, 0].contiguous(), dlt)
        return E.reshape(())

    @staticmethod
    def backward(ctx, gE):
        C, F, V, dlt = ctx.saved_tensors
        jk = ctx.jk
        g = float(gE.item())
        dC = dL = dh = None
        if ctx.needs_input_grad[0]:
            dC = jk.grad_orbitals(C, F, 4.0 * g)
        if ctx.needs_input_grad[1]:
            dL = jk.grad_vectors(C, V, alpha=g, beta=-2.0 * g)
        if ctx.needs_input_grad[2]:
            dh = (g * dlt).reshape(-1)
        return dC, dL, dh, None, None


def closed_shell_energy(C, L, h, jk: JKContext, precision="fp64"):
    """E(C, L, h) with autograd support for C, L and h."""
    return ClosedShellEnergy.apply(C, L, h, jk, precision)
