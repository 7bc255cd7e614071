"""Differentiable far-field spectrum: ``torch.autograd.Function`` over the
HIP forward and adjoint kernels.

    S = far_spectrum(x, y, z, ux, uy, uz, w, layout, grid)      # (1, nPhi, nTheta, nOmega)
    loss = ((S - target)**2 * weights).sum()
    loss.backward()                                              # x.grad ... w.grad

``layout`` is an :class:`engine.HostBatch` giving the CSR offsets, time
alignment and snapshot of the tracks (its own coordinate arrays are not used;
the tensors passed in are).  The |dphase| < pi mask of the forward is held
fixed in the backward; the adjoint recomputes every step tile instead of
storing any intermediate.
"""
from __future__ import annotations

from typing import Optional

import torch

from . import engine


class FarSpectrumFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, z, ux, uy, uz, w, layout: engine.HostBatch, grid: engine.DeviceGrid,
                comp: str, precision: str, tuning: Optional[engine.Tuning]):
        ts = [t.detach().contiguous() for t in (x, y, z, ux, uy, uz, w)]
        batch = engine.DeviceBatch(layout, grid.device, tensors=ts)
        outs = engine.spectrum(batch, grid, comp=comp, precision=precision, tuning=tuning)
        ctx.batch, ctx.grid, ctx.comp, ctx.precision = batch, grid, comp, precision
        return torch.stack(outs)

    @staticmethod
    def backward(ctx, gS):
        dS = [gS[c] for c in range(gS.shape[0])]
        g = engine.far_adjoint(ctx.batch, ctx.grid, dS, comp=ctx.comp, precision=ctx.precision)
        return (*g, None, None, None, None, None)


def far_spectrum(x, y, z, ux, uy, uz, w, layout: engine.HostBatch, grid: engine.DeviceGrid,
                 comp: str = "total", precision: str = "float",
                 tuning: Optional[engine.Tuning] = None) -> torch.Tensor:
    """Far-field spectrum (one snapshot) as a differentiable function of the
    track samples and weights.  Returns (n_out, 1, nPhi, nTheta, nOmega)."""
    if comp not in ("total", "cartesian"):
        raise ValueError("differentiable spectra cover comp 'total' and 'cartesian'")
    if layout.n_snaps != 1:
        raise ValueError("differentiable spectra need exactly one snapshot")
    return FarSpectrumFunction.apply(x, y, z, ux, uy, uz, w, layout, grid, comp, precision, tuning)
