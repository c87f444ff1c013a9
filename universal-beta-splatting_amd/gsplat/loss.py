"""The training step's image loss on the GPU: (1 - lambda) L1 + lambda (1 - SSIM).

What train.py:117-121 computes per view with torch's ``l1_loss``
(utils/loss_utils.py:18-19) and the external ``fused_ssim`` (the SSIM of
utils/loss_utils.py:45-80: 11x11 Gaussian window, sigma 1.5, zero padding,
C1 = 0.01^2, C2 = 0.03^2, mean over every pixel and channel), as one HIP
kernel pair (``ubs_image_loss_fwd/bwd``): the forward also leaves in its
workspace the derivatives the backward needs, so the backward is one more
pass.  Images are [B, H, W, C] or [H, W, C] (the rasteriser's layout); the
loss averages over the whole batch, which for equal-size views is the
reference's mean of per-view losses (train.py:111-128).
"""
from __future__ import annotations

import torch
from torch import Tensor

from .cuda._backend import _F32, lib, ptr, require_cuda, require_dtype, stream_of
from .cuda._wrapper import _call


def _bhwc(t: Tensor) -> Tensor:
    if t.dim() == 3:
        t = t[None]
    assert t.dim() == 4, t.shape
    return t.contiguous()


class _ImageLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, img, gt, lambda_dssim):
        x, y = _bhwc(img), _bhwc(gt.detach())
        assert x.shape == y.shape, (x.shape, y.shape)
        require_cuda(x, y)
        require_dtype(_F32, x, y)
        B, H, W, C = x.shape
        L = lib()
        ws_bytes = int(L.ubs_image_loss_workspace_bytes(B, H, W, C))
        ws = torch.empty((ws_bytes,), dtype=torch.uint8, device=x.device)
        loss = torch.empty((1,), dtype=torch.float32, device=x.device)
        _call("ubs_image_loss_fwd", B, H, W, C, ptr(x), ptr(y), float(lambda_dssim), ptr(loss), ptr(ws), ws_bytes,
              stream_of(x))
        ctx.save_for_backward(x, y, ws)
        ctx.lambda_dssim, ctx.shape = float(lambda_dssim), img.shape
        return loss[0]

    @staticmethod
    def backward(ctx, v_loss):
        x, y, ws = ctx.saved_tensors
        B, H, W, C = x.shape
        v = v_loss.reshape(1).to(torch.float32).contiguous()
        v_img = torch.empty_like(x)
        _call("ubs_image_loss_bwd", B, H, W, C, ptr(x), ptr(y), ctx.lambda_dssim, ptr(v), ptr(ws), ws.numel(),
              ptr(v_img), stream_of(x))
        return v_img.reshape(ctx.shape), None, None


def image_loss(img: Tensor, gt: Tensor, lambda_dssim: float = 0.2) -> Tensor:
    """(1 - lambda_dssim) * L1(img, gt) + lambda_dssim * (1 - SSIM(img, gt)),
    a 0-dim tensor with a gradient for ``img`` (``gt`` gets none, as the
    reference's targets).  lambda_dssim = 0.2 is the reference's default
    (arguments/__init__.py:94)."""
    if not 0.0 <= lambda_dssim <= 1.0:
        raise ValueError("lambda_dssim must be in [0, 1]")
    return _ImageLoss.apply(img, gt, lambda_dssim)


def l1_loss(img: Tensor, gt: Tensor) -> Tensor:
    """mean |img - gt| (utils/loss_utils.py:18-19) in one pass each way."""
    return _ImageLoss.apply(img, gt, 0.0)
