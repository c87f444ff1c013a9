"""Depth -> points / normals for the "Normal" render mode.

Same math as the reference's submodules/gsplat/utils.py:44-131: back-project
each pixel centre with the intrinsics, take central differences of the
resulting points along rows and columns, cross them and normalise; the
one-pixel border is zero.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F
from torch import Tensor


def depth_to_points(depths: Tensor, camtoworlds: Tensor, Ks: Tensor, z_depth: bool = True) -> Tensor:
    """depths [..., H, W, 1] -> world points [..., H, W, 3]."""
    assert depths.shape[-1] == 1, f"Invalid depth shape: {depths.shape}"
    assert camtoworlds.shape[-2:] == (4, 4), f"Invalid viewmats shape: {camtoworlds.shape}"
    assert Ks.shape[-2:] == (3, 3), f"Invalid Ks shape: {Ks.shape}"
    device = depths.device
    height, width = depths.shape[-3:-1]
    x, y = torch.meshgrid(torch.arange(width, device=device), torch.arange(height, device=device),
                          indexing="xy")
    fx, fy = Ks[..., 0, 0], Ks[..., 1, 1]
    cx, cy = Ks[..., 0, 2], Ks[..., 1, 2]
    dirs_cam = torch.stack([
        (x - cx[..., None, None] + 0.5) / fx[..., None, None],
        (y - cy[..., None, None] + 0.5) / fy[..., None, None],
        torch.ones_like(x - cx[..., None, None]),
    ], dim=-1)
    R = camtoworlds[..., :3, :3]
    directions = (R[..., None, None, :, :] @ dirs_cam[..., None]).squeeze(-1)
    origins = camtoworlds[..., :3, -1]
    if not z_depth:
        directions = F.normalize(directions, dim=-1)
    return origins[..., None, None, :] + depths * directions


def depth_to_normal(depths: Tensor, camtoworlds: Tensor, Ks: Tensor, z_depth: bool = True) -> Tensor:
    """depths [..., H, W, 1] -> unit normals [..., H, W, 3] (zero border)."""
    points = depth_to_points(depths, camtoworlds, Ks, z_depth=z_depth)
    dx = points[..., 2:, 1:-1, :] - points[..., :-2, 1:-1, :]
    dy = points[..., 1:-1, 2:, :] - points[..., 1:-1, :-2, :]
    normals = F.normalize(torch.cross(dx, dy, dim=-1), dim=-1)
    return F.pad(normals, (0, 0, 1, 1, 1, 1), value=0.0)
