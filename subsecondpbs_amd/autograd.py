"""Spot-weight gradients through torch.autograd.

``SpotWeightDose.apply(weights, engine, beam)`` returns the beam's dose on the engine's dose
grid as a function of the spot weights (a 1-D float32 tensor, one entry per spot in the beam's
spot order). Its backward is the engine's hand-written adjoint
(``pbs_engine_beam_adjoint``: transposed transfer, transposed superposition gather, the IDD
scaling, then the transposed subspot convolution), evaluated at the forward's weights.

The dose is linear in the weights for fixed live masks, tile radii and step ranges (these do
not depend on the size of the weights), so the gradient is exact up to float rounding wherever
the masks do not change.

``ResidentSpotWeightDose.apply(weights, resident)`` is the same on a resident beam
(``Engine.keep_beam``): the forward is a weight-only update (optionally a HIP graph replay) and
the backward the adjoint from the resident state, so neither repeats the trace or the fill.
"""
from __future__ import annotations

import numpy as np
import torch

from .engine import Engine


class SpotWeightDose(torch.autograd.Function):
    @staticmethod
    def forward(ctx, weights: torch.Tensor, engine: Engine, beam: dict):
        if weights.dim() != 1:
            raise ValueError("weights must be a 1-D tensor, one entry per spot")
        b = dict(beam)
        b["spot_weights"] = weights.detach().to("cpu", torch.float32).numpy().copy()
        nx, ny, nz = engine.dose_dims
        dose = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
        engine.beam_dose(b, dose)
        engine.sync()
        ctx.engine = engine
        ctx.beam = b
        ctx.wdev, ctx.wdtype = weights.device, weights.dtype
        return dose

    @staticmethod
    def backward(ctx, grad_dose: torch.Tensor):
        g = ctx.engine.beam_adjoint(ctx.beam, grad_dose.detach().to(torch.float32).contiguous())
        return torch.from_numpy(np.asarray(g)).to(ctx.wdev, ctx.wdtype), None, None


class ResidentSpotWeightDose(torch.autograd.Function):
    @staticmethod
    def forward(ctx, weights: torch.Tensor, resident):
        if weights.dim() != 1:
            raise ValueError("weights must be a 1-D tensor, one entry per spot")
        w = weights.detach().to("cpu", torch.float32).numpy().copy()
        dose = resident.dose(w)
        ctx.resident, ctx.w = resident, w
        ctx.wdev, ctx.wdtype = weights.device, weights.dtype
        return dose

    @staticmethod
    def backward(ctx, grad_dose: torch.Tensor):
        g = ctx.resident.adjoint(ctx.w, grad_dose.detach().to(torch.float32).contiguous())
        return torch.from_numpy(np.asarray(g)).to(ctx.wdev, ctx.wdtype), None


def beam_dose(weights: torch.Tensor, engine: Engine, beam: dict, resident=None) -> torch.Tensor:
    """Differentiable dose of ``beam`` at ``weights`` (see SpotWeightDose); with ``resident``
    (the beam kept on ``engine``), through ResidentSpotWeightDose."""
    if resident is not None:
        return ResidentSpotWeightDose.apply(weights, resident)
    return SpotWeightDose.apply(weights, engine, beam)


def objective(dose: torch.Tensor, target: torch.Tensor, oar: torch.Tensor, prescription: float = 2.0,
              oar_factor: float = 0.1) -> torch.Tensor:
    """sum over target voxels of (D - prescription)^2 + oar_factor * sum over OAR voxels of D^2."""
    d = dose.to(torch.float64)
    return ((d - prescription) ** 2 * target).sum() + oar_factor * (d ** 2 * oar).sum()
