"""Differentiable far-field spectrum: ``torch.autograd.Function`` over the
HIP forward and adjoint kernels.

    S = far_spectrum(x, y, z, ux, uy, uz, w, layout, grid)      # (1, nPhi, nTheta, nOmega)
    loss = ((S - target)**2 * weights).sum()
    loss.backward()                                              # x.grad ... w.grad

``layout`` is an :class:`engine.HostBatch` giving the CSR offsets, time
alignment and snapshot of the tracks (its own coordinate arrays are not used;
the tensors passed in are).  The |dphase| < pi mask of the forward is held
fixed in the backward.  When every particle's complex field fits
``KEEP_FIELDS_BYTES`` (48 bytes per particle and spectral point), the forward
computes those fields once, forms the spectrum from them and keeps them for
the backward, which then runs only the transpose sweep; otherwise the forward
runs the spectrum kernel and the backward recomputes the fields chunk by
chunk.  No other intermediate is stored.
"""
from __future__ import annotations

from typing import Optional

import torch

from . import engine

# largest per-particle field set the forward keeps for the backward (bytes)
KEEP_FIELDS_BYTES = 8 << 30


class FarSpectrumFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, z, ux, uy, uz, w, layout: engine.HostBatch, grid: engine.DeviceGrid,
                comp: str, precision: str, tuning: Optional[engine.Tuning]):
        ts = [t.detach().contiguous() for t in (x, y, z, ux, uy, uz, w)]
        batch = engine.DeviceBatch(layout, grid.device, tensors=ts)
        ctx.fields = None
        if 0 < engine.field_bytes(batch, grid) <= KEEP_FIELDS_BYTES:
            ctx.fields, outs = engine.far_fields(batch, grid, comp=comp, precision=precision)
        else:
            outs = engine.spectrum(batch, grid, comp=comp, precision=precision, tuning=tuning)
        ctx.batch, ctx.grid, ctx.comp, ctx.precision = batch, grid, comp, precision
        return torch.stack(outs)

    @staticmethod
    def backward(ctx, gS):
        dS = [gS[c] for c in range(gS.shape[0])]
        g = engine.far_adjoint(ctx.batch, ctx.grid, dS, comp=ctx.comp, precision=ctx.precision,
                               fields=ctx.fields)
        ctx.fields = None
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
