"""Drop-in stand-in for the reference's compiled extension module `_C`.

The reference's submodules/gsplat/cuda/_wrapper.py reaches its kernels as
``_make_lazy_cuda_func(name)(*args)``, i.e. ``getattr(_C, name)(*args)``
with ``_C`` built from cuda/csrc/ext.cpp (:3-58).  This module exposes the
functions of that module the rasteriser path calls, with the same names,
positional arguments and return tuples (cuda/csrc/bindings.h), implemented
on libubs_hip.so.  A maintainer who keeps the reference's own _wrapper.py
points its ``from ._backend import _C`` at this module instead.

Outside the path (not provided): quat_scale_to_covar_preci_*, proj_*,
world_to_cam_*, rasterize_to_indices_in_range, and the packed (camera_ids /
primitive_ids) form of isect_tiles.
"""
from __future__ import annotations

from . import _wrapper as W


# ---- conditioning and covariance (ext.cpp:5-24)
def cond_mean_convariance_opacity_fwd(means, covars, opacities, betas, query):
    return W.cond_fwd_raw(means.contiguous(), covars.contiguous(), opacities.contiguous(), betas.contiguous(),
                          query.contiguous())


def cond_mean_convariance_opacity_bwd(means, covars, opacities, betas, query, v_means, v_covars, v_opacities):
    return W.cond_bwd_raw(means, covars, opacities, betas, query, v_means, v_covars, v_opacities)


def rot_scale_l_triangle_to_covar_fwd(rot, scale, l_triangle, rest_i, rest_j, spatial_block):
    return W.covar_fwd_raw(rot.contiguous(), scale.contiguous(), l_triangle.contiguous(), rest_i, rest_j,
                           spatial_block)


def rot_scale_l_triangle_to_covar_bwd(rot, scale, l_triangle, rest_i, rest_j, spatial_block, v_covar):
    return W.covar_bwd_raw(rot, scale, l_triangle, rest_i, rest_j, spatial_block, v_covar)


def l_triangle_to_rotmat_fwd(l_triangle):
    return W.rotmat_fwd_raw(l_triangle.contiguous())


def l_triangle_to_rotmat_bwd(l_triangle, v_rots):
    return W.rotmat_bwd_raw(l_triangle, v_rots)


# ---- projection (ext.cpp:41-46)
def fully_fused_projection_fwd(means, covars, quats, scales, viewmats, Ks, image_width, image_height, eps2d,
                               near_plane, far_plane, radius_clip, calc_compensations, ortho):
    return W.projection_fwd_raw(means, covars, quats, scales, viewmats, Ks, image_width, image_height, eps2d,
                                near_plane, far_plane, radius_clip, calc_compensations, ortho)


def fully_fused_projection_bwd(means, covars, quats, scales, viewmats, Ks, image_width, image_height, eps2d,
                               ortho, radii, conics, compensations, v_means2d, v_depths, v_conics,
                               v_compensations, viewmats_requires_grad):
    return W.projection_bwd_raw(means, covars, quats, scales, viewmats, Ks, image_width, image_height, eps2d,
                                ortho, radii, conics, compensations, v_means2d, v_depths, v_conics,
                                v_compensations, viewmats_requires_grad)


# ---- intersections (ext.cpp:48-49)
def isect_tiles(means2d, radii, depths, camera_ids, primitive_ids, C, tile_size, tile_width, tile_height, sort,
                double_buffer):
    if camera_ids is not None or primitive_ids is not None:
        raise NotImplementedError("packed isect_tiles is outside the rasteriser path")
    return W.isect_tiles(means2d, radii, depths, tile_size, tile_width, tile_height, sort)


def isect_offset_encode(isect_ids, C, tile_width, tile_height):
    return W.isect_offset_encode(isect_ids, C, tile_width, tile_height)


# ---- blend (ext.cpp:51-52); the reference's backward re-derives T from alpha
def rasterize_to_pixels_fwd(means2d, conics, colors, opacities, betas, backgrounds, masks, image_width,
                            image_height, tile_size, tile_offsets, flatten_ids):
    rc, ra, last, _ = W.raster_fwd_raw(means2d, conics, colors, opacities, betas, backgrounds, masks, image_width,
                                       image_height, tile_size, tile_offsets, flatten_ids)
    return rc, ra, last


def rasterize_to_pixels_bwd(means2d, conics, colors, opacities, betas, backgrounds, masks, image_width,
                            image_height, tile_size, tile_offsets, flatten_ids, render_alphas, last_ids,
                            v_render_colors, v_render_alphas):
    return W.raster_bwd_raw(means2d, conics, colors, opacities, betas, backgrounds, masks, image_width,
                            image_height, tile_size, tile_offsets, flatten_ids, render_alphas, last_ids, None,
                            v_render_colors, v_render_alphas)
