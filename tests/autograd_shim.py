"""torch.autograd.Function over the C-ABI (ctypes), for the test harness only.

Filtered backprojection  img = BP(F_h(sino))  with the hand-written adjoint:
  d sino = F_h^T BP^T d img      (the sheared-path projector, then the conjugate-window filter)
  d h    = grad_taps(sino, BP^T d img)
Only the unfiltered input is kept for the backward pass; the filtered sinogram
and the image are recomputed-free (the operators are linear), so nothing of
the forward's intermediates is stored.
"""
from __future__ import annotations

import torch


class FilteredBackprojection(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, taps, filt, bp):
        S, P, D = sino.shape
        N = bp.g.numberOfPixels
        filt.set_taps(taps.detach().float().cpu().numpy())
        x = sino.detach().contiguous().clone()
        filt(x, S)
        img = torch.empty((S, N, N), dtype=torch.float32, device=sino.device)
        bp.batch(x, S, img)
        ctx.save_for_backward(sino.detach())
        ctx.filt, ctx.bp, ctx.taps_device = filt, bp, taps.device
        return img

    @staticmethod
    def backward(ctx, gimg):
        (sino,) = ctx.saved_tensors
        filt, bp = ctx.filt, ctx.bp
        S, P, D = sino.shape
        gimg = gimg.contiguous()
        gs = torch.empty((S, P, D), dtype=torch.float32, device=sino.device)
        bp.adjoint_batch(gimg, S, gs)
        dtaps = torch.empty(D // 2 + 1, dtype=torch.float32, device=sino.device)
        filt.grad_taps(sino.contiguous(), gs, S, dtaps)
        filt.adjoint(gs, S)
        return gs, dtaps.to(ctx.taps_device), None, None


def fbp(sino, taps, filt, bp):
    return FilteredBackprojection.apply(sino, taps, filt, bp)
