"""Pure-torch restatements of the covariance and conditioning ops.

The reference model imports these names (scene/beta_model.py:13-17) from its
submodules/gsplat/cuda/_torch_impl.py but never calls them on the render
path; they are provided so that import keeps working.  They run on any
device through plain torch ops and are not used by the HIP path.
"""
from __future__ import annotations

from typing import Tuple

import torch
from torch import Tensor


def _cond_mean_convariance_opacity(means: Tensor, covars: Tensor, opacities: Tensor, betas: Tensor,
                                   query: Tensor) -> Tuple[Tensor, Tensor, Tensor]:
    """Gaussian conditioning with per-dimension Beta control (see
    cond_mean_convariance_opacity); returns m [N,3], V [N,3,3], o [N,1]."""
    v11, v12 = covars[:, :3, :3], covars[:, :3, 3:]
    v21, v22 = covars[:, 3:, :3], covars[:, 3:, 3:]
    x = query - means[:, 3:]
    regr = v12 @ torch.linalg.inv(v22)
    badj = torch.clamp_max(betas / 4.0, 1.0)
    rb = regr * badj[:, None, :]
    m = means[:, :3] + (rb @ x[..., None])[..., 0]
    V = v11 - rb @ v21
    L = torch.linalg.cholesky(v22)
    y = torch.linalg.solve_triangular(L, x[..., None], upper=False)[..., 0]
    d = torch.tanh(y * y).clamp(0.0, 1.0 - torch.finfo(covars.dtype).eps)
    o = opacities * torch.prod((1.0 - d) ** betas, dim=-1, keepdim=True)
    return m, V, o


def _l_triangle_to_rotmat(l_triangle: Tensor) -> Tensor:
    """R = I + A with A skew from l_triangle[:, :3] = (a01, a02, a12)."""
    a01, a02, a12 = l_triangle[:, 0], l_triangle[:, 1], l_triangle[:, 2]
    one = torch.ones_like(a01)
    R = torch.stack([one, a01, a02, -a01, one, a12, -a02, -a12, one], dim=-1)
    return R.reshape(-1, 3, 3)


def _rot_scale_l_triangle_to_covar(rot: Tensor, scale: Tensor, l_triangle: Tensor, rest_i: Tensor,
                                   rest_j: Tensor, spatial_block: bool = False) -> Tensor:
    """Sigma = L L^T with L[:3,:3] = rot diag(scale[:3]), L[i,i] = scale[i]
    (i >= 3) and L[rest_i, rest_j] = l_triangle[:, 3:]."""
    Lxyz = rot * scale[:, None, :3]
    if spatial_block:
        return Lxyz @ Lxyz.transpose(-1, -2)
    N, D = scale.shape
    L = scale.new_zeros((N, D, D))
    L[:, :3, :3] = Lxyz
    idx = torch.arange(3, D, device=scale.device)
    L[:, idx, idx] = scale[:, 3:]
    L[:, rest_i.long(), rest_j.long()] = l_triangle[:, 3:3 + rest_i.numel()]
    return L @ L.transpose(-1, -2)
