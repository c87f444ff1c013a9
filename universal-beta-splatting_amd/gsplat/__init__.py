"""Drop-in gsplat package for universal-beta-splatting on AMD MI355X (gfx950).

Same import surface as the reference's submodules/gsplat for the Beta-splat
rasteriser path: ``gsplat.rendering.rasterization`` and the ops in
``gsplat.cuda._wrapper``, backed by hand-written HIP kernels in libubs_hip.so.
"""
from .cuda._wrapper import (
    cond_mean_convariance_opacity,
    fully_fused_projection,
    isect_offset_encode,
    isect_tiles,
    l_triangle_to_rotmat,
    rasterize_to_pixels,
    rot_scale_l_triangle_to_covar,
)
from .rendering import rasterization
from .version import __version__

__all__ = [
    "rasterization",
    "cond_mean_convariance_opacity",
    "fully_fused_projection",
    "isect_offset_encode",
    "isect_tiles",
    "l_triangle_to_rotmat",
    "rasterize_to_pixels",
    "rot_scale_l_triangle_to_covar",
    "__version__",
]
