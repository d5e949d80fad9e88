"""Differentiable line-of-sight far field for gradient-based beam synthesis.

`FarField.apply(w, pos, sinks, sink_normals, hit_bits, alpha, beta)` returns the
source-summed field E (K, 3) complex64 of `lem_field_los` and differentiates it
with the hand-written adjoint kernel `lem_field_los_adjoint`: gradients flow to
the complex excitation weights w (S, 3) and the source positions pos (S, 3);
the visibility mask (hit_bits) is frozen.  The forward arithmetic is the
reference's polarised Rayleigh-Sommerfeld path term (empropagation.py:768-801,
178-249); the reference itself has no gradients.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native


class FarField(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w, pos, sinks, sink_normals, hit_bits, alpha, beta):
        E = _native.field_los(pos.detach(), w.detach(), sinks, sink_normals, hit_bits=hit_bits,
                              alpha=alpha, beta=beta)
        ctx.save_for_backward(w.detach(), pos.detach(), sinks, sink_normals)
        ctx.hit_bits = hit_bits
        ctx.alpha, ctx.beta = alpha, beta
        return E

    @staticmethod
    def backward(ctx, grad_E):
        w, pos, sinks, sink_normals = ctx.saved_tensors
        gw, gp = _native.field_los_adjoint(
            pos, w, sinks, sink_normals, grad_E.contiguous(), hit_bits=ctx.hit_bits,
            alpha=ctx.alpha, beta=ctx.beta, want_w=ctx.needs_input_grad[0],
            want_p=ctx.needs_input_grad[1])
        return gw, gp, None, None, None, None, None


def far_field(w, pos, sinks, sink_normals, hit_bits=None, alpha=0.0, beta=2 * np.pi):
    return FarField.apply(w, pos, sinks, sink_normals, hit_bits, float(alpha), float(beta))


def db_mask_loss(E, in_beam, sidelobe_db=-25.0, ref_power=None, eps=1e-30):
    """Squared dB mismatch to a mask: 0 dB inside the main beam, at most
    `sidelobe_db` outside (excess only), relative to `ref_power` (default: the
    current peak power, treated as a constant).  Mean over sinks."""
    P = (E.real ** 2 + E.imag ** 2).sum(dim=-1)
    peak = P.detach().max() if ref_power is None else ref_power
    db = 10.0 * torch.log10(P / peak + eps)
    err = torch.where(in_beam, db, torch.clamp(db - sidelobe_db, min=0.0))
    return (err ** 2).mean()
