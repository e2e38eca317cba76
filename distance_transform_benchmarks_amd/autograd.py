"""Differentiable GGDT: torch.autograd.Function over the HIP forward and adjoint.

forward : converged GGDT (task schedule) of float32 images with soft float
          masks, then the argmin predecessor codes (saved for backward).
backward: the adjoint kernels of libdtk (dtk_ggdt_backward_batched, one
          launch for the whole batch) scatter dL/dD along the predecessor
          forest into dL/dimage and dL/dmask.
Distances stay float32; rounding is treated as exact in the adjoint.
No reference counterpart (the reference GGDT is forward-only).
"""
from __future__ import annotations

import torch

from . import api


class GGDTFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, img, mask, lambda_: float, nu: float, connectivity: int = 8):
        if img.dtype != torch.float32 or mask.dtype != torch.float32:
            raise ValueError("the differentiable GGDT takes float32 images and soft float32 masks")
        squeeze = img.dim() == 2
        imgs = img.unsqueeze(0) if squeeze else img
        masks = mask.unsqueeze(0) if squeeze else mask
        imgs = imgs.contiguous()
        masks = masks.contiguous()
        D = torch.empty(imgs.shape, dtype=torch.float32, device=img.device)
        P = torch.empty(imgs.shape, dtype=torch.uint8, device=img.device)
        if imgs.shape[0] > 1:
            api.ggdt_batched(list(imgs), list(masks), lambda_, nu, connectivity=connectivity, outs=list(D))
        elif imgs.shape[0] == 1:
            api.ggdt(imgs[0], masks[0], lambda_, nu, connectivity=connectivity, out=D[0])
        if imgs.shape[0] > 1:
            api.predecessors_batched(list(imgs), list(masks), list(D), lambda_, nu, connectivity=connectivity,
                                     outs=list(P))
        elif imgs.shape[0] == 1:
            api.predecessors(imgs[0], masks[0], D[0], lambda_, nu, connectivity=connectivity, out=P[0])
        ctx.save_for_backward(imgs, masks, D, P)
        ctx.lambda_, ctx.nu, ctx.connectivity, ctx.squeeze = lambda_, nu, connectivity, squeeze
        return D[0] if squeeze else D

    @staticmethod
    def backward(ctx, grad_D):
        imgs, masks, D, P = ctx.saved_tensors
        g = grad_D.unsqueeze(0) if ctx.squeeze else grad_D
        g = g.contiguous().float()
        gis, gms = api.backward_batched(list(imgs), list(masks), list(D), list(P), list(g), ctx.lambda_, ctx.nu,
                                        connectivity=ctx.connectivity)
        gi = torch.stack(gis) if gis else torch.empty_like(imgs)
        gm = torch.stack(gms) if gms else torch.empty_like(masks)
        if ctx.squeeze:
            gi, gm = gi[0], gm[0]
        return gi, gm, None, None, None


def ggdt_differentiable(img, mask, lambda_: float = 1.0, nu: float = 1e3, connectivity: int = 8):
    """D = GGDT(img, soft mask) with gradients w.r.t. img and mask."""
    return GGDTFunction.apply(img, mask, lambda_, nu, connectivity)
