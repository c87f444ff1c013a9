"""Beta-model glue around the rasteriser: parameter activations, view/time
query, conditioning and the rasterization() call (render), and the fused
multi-view path (render_views) that does all of it but the blend in one
kernel per direction.

Mirrors BetaModel.render (scene/beta_model.py:660-722) and its getters
(:36-52, :103-159): scale = softplus, opacity = sigmoid, beta = 4 exp(b),
mean = cat(xyz, mean_rest), query = normalise(xyz - camera centre) (+ t for
7-D), conditioned by cond_mean_convariance_opacity and rendered with the
conditioned 3x3 covariances.  The query carries no gradient into the
conditioning op, as in the reference.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, Optional

import torch
import torch.nn.functional as F
from torch import Tensor

from .cuda._wrapper import (
    beta_activations,
    beta_project,
    camera_centers,
    prepacked_blend_workspace,
    cond_mean_convariance_opacity,
    isect_begin,
    isect_finish,
    isect_read_info,
    isect_offset_encode,
    isect_tiles_offsets,
    l_triangle_to_rotmat,
    rasterize_to_pixels,
    rot_scale_l_triangle_to_covar,
)
from .rendering import rasterization


def quat_to_rotmat(q: Tensor) -> Tensor:
    """Differentiable rotation from (w, x, y, z) quaternions (normalised here)."""
    q = F.normalize(q, dim=-1)
    w, x, y, z = q.unbind(-1)
    R = torch.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
        2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
        2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y),
    ], dim=-1)
    return R.reshape(q.shape[:-1] + (3, 3))


@dataclass
class BetaParams:
    """Leaf parameters of a D-dim Beta scene, pre-activation (BetaModel layout)."""
    xyz: Tensor          # [N,3]
    mean_rest: Tensor    # [N,D-3]
    scale_raw: Tensor    # [N,D]
    l_triangle: Tensor   # [N,D(D-1)/2]
    opacity_raw: Tensor  # [N,1]
    beta_raw: Tensor     # [N,D-2]
    rgb: Tensor          # [N,3]
    quats: Optional[Tensor] = None  # [N,4]: rotation from quaternions instead of l_triangle[:, :3]

    @property
    def D(self) -> int:
        return self.scale_raw.shape[1]

    def leaves(self) -> Dict[str, Tensor]:
        d = dict(xyz=self.xyz, mean_rest=self.mean_rest, scale_raw=self.scale_raw,
                 l_triangle=self.l_triangle, opacity_raw=self.opacity_raw, beta_raw=self.beta_raw,
                 rgb=self.rgb)
        if self.quats is not None:
            d["quats"] = self.quats
        return d

    @staticmethod
    def from_ply(path: str, device, requires_grad: bool = True) -> "BetaParams":
        """A scene saved by BetaModel.save_ply (scene/beta_model.py:298-320),
        resident on `device` as one tensor per leaf (rotation from l_triangle)."""
        from .ply import load_ply

        d = load_ply(path)

        def t(a):
            return torch.tensor(a, dtype=torch.float32, device=device).requires_grad_(requires_grad)
        return BetaParams(xyz=t(d["xyz"]), mean_rest=t(d["mean_rest"]), scale_raw=t(d["scale_raw"]),
                          l_triangle=t(d["l_triangle"]), opacity_raw=t(d["opacity_raw"]), beta_raw=t(d["beta_raw"]),
                          rgb=t(d["rgb"]))

    def save_ply(self, path: str) -> None:
        """Write the leaves in the reference's PLY layout (quaternions, which
        that layout has no property for, are not stored)."""
        from .ply import save_ply

        save_ply(path, self.xyz, self.mean_rest, self.scale_raw, self.l_triangle, self.opacity_raw, self.beta_raw,
                 self.rgb)

    @staticmethod
    def from_scene(sc, device, requires_grad: bool = True, use_quats: bool = True) -> "BetaParams":
        def t(a):
            return torch.tensor(a, dtype=torch.float32, device=device).requires_grad_(requires_grad)
        return BetaParams(xyz=t(sc.xyz), mean_rest=t(sc.mean_rest), scale_raw=t(sc.scale_raw),
                          l_triangle=t(sc.l_triangle), opacity_raw=t(sc.opacity_raw),
                          beta_raw=t(sc.beta_raw), rgb=t(sc.rgb),
                          quats=t(sc.quats) if use_quats else None)


def rest_indices(D: int, device) -> tuple:
    """rest_i, rest_j of scene/beta_model.py:69-73 as device int32 tensors."""
    ti, tj = torch.tril_indices(D, D, offset=-1)
    m = (ti >= 3) | (tj >= 3)
    return ti[m].to(torch.int32).to(device), tj[m].to(torch.int32).to(device)


def camera_center(viewmat: Tensor) -> Tensor:
    """World position of a world->camera [4,4] row-major matrix (the bits
    the fused projection takes it with)."""
    return camera_centers(viewmat[None])[0]


def conditioned(params: BetaParams, viewmat: Tensor, timestamp: Optional[float] = None, rest=None):
    """Activations + query + conditioning: (means [N,3], covars [N,3,3],
    opacities [N], beta_xyz [N], scale [N,D]).

    The activations, the query and the quaternion rotation run as one kernel
    (beta_activations); the rotation from l_triangle, the covariance and the
    conditioning are the reference's ops."""
    D = params.D
    if D not in (6, 7):
        raise NotImplementedError("Only implemented for 6D or 7D query")
    with torch.no_grad():
        cam_pos = camera_center(viewmat).to(torch.float32).contiguous()
    mean, scale, opac, beta, query, rot = beta_activations(
        params.xyz, params.mean_rest, params.scale_raw, params.opacity_raw, params.beta_raw, params.quats,
        cam_pos, float(timestamp or 0.0))
    if rot is None:
        rot = l_triangle_to_rotmat(params.l_triangle[:, :3])
    ri, rj = rest if rest is not None else rest_indices(D, params.xyz.device)
    cov = rot_scale_l_triangle_to_covar(rot, scale, params.l_triangle, ri, rj)
    m, V, o = cond_mean_convariance_opacity(mean, cov, opac, beta[:, 1:], query)
    return m, V, o.squeeze(-1), beta[:, 0], scale


def render(params: BetaParams, viewmat: Tensor, K: Tensor, width: int, height: int,
           timestamp: Optional[float] = None, background: Optional[Tensor] = None, rest=None,
           **raster_kwargs):
    """One view through the drop-in path (BetaModel.render): returns
    (render_colors [1,H,W,3], render_alphas [1,H,W,1], meta)."""
    m, V, o, beta0, scale = conditioned(params, viewmat, timestamp, rest)
    bg = background[None] if background is not None else None
    return rasterization(
        means=m, l_triagnles=params.l_triangle, scales=scale, opacities=o, betas=beta0,
        colors=params.rgb, viewmats=viewmat[None], Ks=K[None], width=width, height=height,
        backgrounds=bg, covars=V, **raster_kwargs)


def render_views(params: BetaParams, viewmats: Tensor, Ks: Tensor, width: int, height: int,
                 timestamps: Optional[Tensor] = None, backgrounds: Optional[Tensor] = None,
                 tile_size: int = 16, near_plane: float = 0.01, far_plane: float = 1e10,
                 radius_clip: float = 0.0, eps2d: float = 0.3, cam_pos: Optional[Tensor] = None):
    """C views of one scene in one pass: the fused Beta projection
    (beta_project: activations, query, covariance, per-view conditioning and
    projection in one launch), then the reference's intersection, sort and
    blend over all C views.  Returns (render_colors [C,H,W,3],
    render_alphas [C,H,W,1], meta) like rasterization(); with one view the
    images and every leaf gradient equal render()'s bit for bit.

    Each view is conditioned on its own query, which is what BetaModel.render
    does one view at a time (scene/beta_model.py:660-722); gradients of all
    views add up in the leaves.  cam_pos [C,3] are the camera centres the
    queries are taken from (default: derived from viewmats)."""
    D = params.D
    if D not in (6, 7):
        raise NotImplementedError("Only implemented for 6D or 7D query")
    C = viewmats.shape[0]
    assert viewmats.shape == (C, 4, 4) and Ks.shape == (C, 3, 3), (viewmats.shape, Ks.shape)
    if cam_pos is not None:  # (None: the projection takes each view's centre itself)
        cam_pos = cam_pos.detach().to(torch.float32).contiguous()
    tw = math.ceil(width / float(tile_size))
    th = math.ceil(height / float(tile_size))
    # the projection also writes the blend's per-entry records (no packing pass)
    ws = prepacked_blend_workspace(C, params.xyz.shape[0], params.rgb.shape[-1], tile_size, tw, th,
                                   params.xyz.device)
    radii, means2d, depths, conics, opacities, betas = beta_project(
        params.xyz, params.mean_rest, params.scale_raw, params.l_triangle, params.opacity_raw, params.beta_raw,
        params.quats, viewmats, Ks, width, height, cam_pos=cam_pos, timestamps=timestamps, eps2d=eps2d,
        near_plane=near_plane, far_plane=far_plane, radius_clip=radius_clip, rgb=params.rgb, records=ws)
    tiles_per_gauss, isect_ids, flatten_ids, isect_offsets = isect_tiles_offsets(means2d, radii, depths, tile_size,
                                                                                tw, th)
    if isect_offsets is None:
        isect_offsets = isect_offset_encode(isect_ids, C, tw, th)
    # (one view: a view of the leaf, whose backward is no kernel; several: their gradients sum)
    colors = params.rgb[None] if C == 1 else params.rgb.expand(C, -1, -1)
    render_colors, render_alphas = rasterize_to_pixels(
        means2d, conics, colors, opacities, betas, width, height, tile_size, isect_offsets, flatten_ids,
        backgrounds=backgrounds, prepacked_workspace=ws)
    meta = dict(radii=radii, means2d=means2d, depths=depths, conics=conics, opacities=opacities, betas=betas,
                tile_width=tw, tile_height=th, tiles_per_gauss=tiles_per_gauss, isect_ids=isect_ids,
                flatten_ids=flatten_ids, isect_offsets=isect_offsets, width=width, height=height,
                tile_size=tile_size, n_cameras=C)
    return render_colors, render_alphas, meta


def render_batch(params: BetaParams, viewmats: Tensor, Ks: Tensor, sizes, timestamps: Optional[Tensor] = None,
                 backgrounds: Optional[Tensor] = None, tile_size: int = 16, near_plane: float = 0.01,
                 far_plane: float = 1e10, radius_clip: float = 0.0, eps2d: float = 0.3,
                 cam_pos: Optional[Tensor] = None):
    """A batch of C views of one scene at per-view image sizes (sizes: C
    pairs (W, H)), e.g. one training step over several cameras: one fused
    Beta projection for all views (per-view conditioning and image size),
    then intersection, sort and blend per group of views sharing a size.
    Returns (list of render_colors [H_c,W_c,3], list of render_alphas
    [H_c,W_c,1]) in view order; gradients of every view add up in the
    leaves.  The reference renders such a batch one view at a time
    (train.py:111-128, BetaModel.render per view)."""
    D = params.D
    if D not in (6, 7):
        raise NotImplementedError("Only implemented for 6D or 7D query")
    C = viewmats.shape[0]
    sizes = [(int(w), int(h)) for w, h in sizes]
    assert len(sizes) == C, (len(sizes), C)
    dev = viewmats.device
    if cam_pos is not None:  # (None: the projection takes each view's centre itself)
        cam_pos = cam_pos.detach().to(torch.float32).contiguous()
    # project the views ordered by size group, so that each group's
    # projected rows are one contiguous slice [c0:c1] of the [C,N,*] outputs
    order = sorted(range(C), key=lambda c: (sizes[c], c))
    if order != list(range(C)):
        perm = torch.tensor(order, dtype=torch.long, device=dev)
        viewmats, Ks = viewmats.index_select(0, perm), Ks.index_select(0, perm)
        if cam_pos is not None:
            cam_pos = cam_pos.index_select(0, perm)
        if timestamps is not None:
            timestamps = timestamps.index_select(0, perm)
        if backgrounds is not None:
            backgrounds = backgrounds.index_select(0, perm)
    wh = torch.tensor([sizes[c] for c in order], dtype=torch.int32, device=dev)
    radii, means2d, depths, conics, opacities, betas = beta_project(
        params.xyz, params.mean_rest, params.scale_raw, params.l_triangle, params.opacity_raw, params.beta_raw,
        params.quats, viewmats, Ks, 0, 0, cam_pos=cam_pos, timestamps=timestamps, image_sizes=wh, eps2d=eps2d,
        near_plane=near_plane, far_plane=far_plane, radius_clip=radius_clip)
    colors_out, alphas_out = [None] * C, [None] * C
    groups, c0 = [], 0
    while c0 < C:
        c1 = c0
        while c1 < C and sizes[order[c1]] == sizes[order[c0]]:
            c1 += 1
        groups.append((c0, c1))
        c0 = c1
    # every group's counts first, then one host read of all their n_isects
    states = []
    for c0, c1 in groups:
        W, H = sizes[order[c0]]
        tw, th = math.ceil(W / float(tile_size)), math.ceil(H / float(tile_size))
        states.append(isect_begin(means2d[c0:c1], radii[c0:c1], depths[c0:c1], tile_size, tw, th))
    infos = isect_read_info(states)
    for g, (c0, c1) in enumerate(groups):
        W, H = sizes[order[c0]]
        Cg = c1 - c0
        tw, th = math.ceil(W / float(tile_size)), math.ceil(H / float(tile_size))
        _, isect_ids, flatten_ids, offsets = isect_finish(states[g], *infos[g])
        if offsets is None:
            offsets = isect_offset_encode(isect_ids, Cg, tw, th)
        bg = backgrounds[c0:c1] if backgrounds is not None else None
        rc, ra = rasterize_to_pixels(means2d[c0:c1], conics[c0:c1], params.rgb.expand(Cg, -1, -1), opacities[c0:c1],
                                     betas[c0:c1], W, H, tile_size, offsets, flatten_ids, backgrounds=bg)
        for k in range(Cg):
            colors_out[order[c0 + k]], alphas_out[order[c0 + k]] = rc[k], ra[k]
    return colors_out, alphas_out
