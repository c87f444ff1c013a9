"""Loader of libubs_hip.so, the gfx950 HIP library behind this package.

Replaces the reference's extension loader (submodules/gsplat/cuda/_backend.py:79-141):
instead of a pybind11 torch extension, the kernels sit behind the plain C ABI
declared in include/ubs_capi.h and are called through ctypes with raw device
pointers and the current HIP stream.  There is no fallback: if the library or
a GPU is missing, every call raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("UBS_HIP_LIB", os.path.join(_HERE, "libubs_hip.so"))

_c_i32 = ctypes.c_int32
_c_u32 = ctypes.c_uint32
_c_i64 = ctypes.c_int64
_c_f32 = ctypes.c_float
_c_int = ctypes.c_int
_c_size = ctypes.c_size_t
_vp = ctypes.c_void_p

# name -> (restype, argtypes)
_SIGNATURES = {
    "ubs_version": (_c_int, []),
    "ubs_last_error": (ctypes.c_char_p, []),
    "ubs_beta_activate_fwd": (
        _c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_f32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ubs_beta_activate_bwd": (
        _c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ubs_l_triangle_to_rotmat_fwd": (_c_int, [_c_i64, _vp, _vp, _vp]),
    "ubs_l_triangle_to_rotmat_bwd": (_c_int, [_c_i64, _vp, _vp, _vp]),
    "ubs_rot_scale_l_triangle_to_covar_fwd": (
        _c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _vp, _vp, _c_i64, _c_int, _vp, _vp]),
    "ubs_rot_scale_l_triangle_to_covar_bwd": (
        _c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _vp, _vp, _c_i64, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "ubs_cond_mean_convariance_opacity_fwd": (
        _c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ubs_cond_mean_convariance_opacity_bwd": (
        _c_int, [_c_i64, _c_i64] + [_vp] * 12 + [_vp]),
    "ubs_fully_fused_projection_fwd": (
        _c_int, [_c_u32, _c_u32, _vp, _vp, _c_int, _vp, _vp, _vp, _vp, _c_i32, _c_i32, _c_f32, _c_f32, _c_f32,
                 _c_f32, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ubs_fully_fused_projection_bwd": (
        _c_int, [_c_u32, _c_u32, _vp, _vp, _c_int, _vp, _vp, _vp, _vp, _c_i32, _c_i32, _c_f32, _c_int, _vp, _vp,
                 _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ubs_beta_project_fwd": (
        _c_int, [_c_i64, _c_i64, _c_u32] + [_vp] * 11 + [_c_i32, _c_i32, _vp, _c_f32, _c_f32, _c_f32, _c_f32,
                                                       _c_int] + [_vp] * 7),
    "ubs_beta_project_fwd_rec": (
        _c_int, [_c_i64, _c_i64, _c_u32] + [_vp] * 11 + [_c_i32, _c_i32, _vp, _c_f32, _c_f32, _c_f32, _c_f32,
                                                       _c_int] + [_vp] * 9),
    "ubs_beta_project_bwd": (
        _c_int, [_c_i64, _c_i64, _c_u32] + [_vp] * 11 + [_c_i32, _c_i32, _vp, _c_f32, _c_int] + [_vp] * 15),
    "ubs_camera_centers": (_c_int, [_c_u32, _vp, _vp, _vp]),
    "ubs_isect_count_workspace_bytes": (_c_size, [_c_u32, _c_u32]),
    "ubs_isect_tiles_count": (
        _c_int, [_c_u32, _c_u32, _vp, _vp, _c_u32, _c_u32, _c_u32, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "ubs_isect_tiles_emit": (
        _c_int, [_c_u32, _c_u32, _vp, _vp, _vp, _vp, _c_u32, _c_u32, _c_u32, _vp, _vp, _vp]),
    "ubs_isect_prefix_bits": (_c_u32, [_c_u32, _c_u32, _c_u32]),
    "ubs_radix_sort_workspace_bytes": (_c_size, [_c_i64, _c_int, _c_int]),
    "ubs_radix_sort_pairs": (
        _c_int, [_c_i64, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp, _c_size, ctypes.POINTER(_c_int), _vp]),
    "ubs_radix_sort_pairs_flag": (
        _c_int, [_c_i64, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_u32, _vp, _c_size, ctypes.POINTER(_c_int), _vp, _vp]),
    "ubs_isect_tiles_count_info": (
        _c_int, [_c_u32, _c_u32, _vp, _vp, _vp, _c_u32, _c_u32, _c_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                 _c_size, _vp]),
    "ubs_isect_tiles_emit_ordered_rec": (
        _c_int, [_c_u32, _c_u32, _vp, _vp, _vp, _c_u32, _c_u32, _c_u32, _vp, _vp, _c_i64, _vp]),
    "ubs_radix_sort_pairs_skip": (
        _c_int, [_c_i64, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_u32, _vp, _c_size, ctypes.POINTER(_c_int), _vp]),
    "ubs_beta_project_fwd_rows": (
        _c_int, [_c_i64, _c_i64, _c_u32] + [_vp] * 11 + [_c_i32, _c_i32, _vp, _c_f32, _c_f32, _c_f32, _c_f32,
                                                       _c_int] + [_vp] * 6 + [_c_i64, _vp]),
    "ubs_beta_project_fwd_table": (
        _c_int, [_c_u32, _c_i64] + [_vp] * 6 + [_c_f32, _c_f32, _c_f32, _c_f32] + [_vp] * 7 + [_vp]),
    "ubs_isect_tiles_emit_ordered_cap": (
        _c_int, [_c_u32, _c_u32, _vp, _vp, _vp, _vp, _vp, _c_u32, _c_u32, _c_u32, _vp, _vp, _c_i64, _vp]),
    "ubs_radix_sort_pairs_dev": (
        _c_int, [_c_i64, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp, _c_size, ctypes.POINTER(_c_int), _vp]),
    "ubs_isect_offset_encode_dev": (_c_int, [_c_i64, _vp, _vp, _c_u32, _c_u32, _c_u32, _vp, _vp]),
    "ubs_radix_sort_pairs_noalloc": (
        _c_int, [_c_i64, _vp, _vp, _vp, _vp, _c_int, _c_int, ctypes.POINTER(_c_int), _vp, _vp]),
    "ubs_radix_sort_status": (_c_int, [_vp, ctypes.POINTER(_c_int), _vp]),
    "ubs_isect_depth_keys": (_c_int, [_c_u32, _c_u32, _vp, _vp, _vp, _vp, _vp]),
    "ubs_isect_ordered_workspace_bytes": (_c_size, [_c_u32, _c_u32]),
    "ubs_isect_scan_ordered": (_c_int, [_c_u32, _c_u32, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "ubs_isect_tiles_emit_ordered": (
        _c_int, [_c_u32, _c_u32, _vp, _vp, _vp, _vp, _vp, _c_u32, _c_u32, _c_u32, _vp, _vp, _vp]),
    "ubs_sort_tile_segments": (_c_int, [_c_i64, _vp, _vp, _vp, _vp, _vp, _c_u32, _c_u32, _c_u32, _c_u32, _vp]),
    "ubs_isect_tiles_count_range": (
        _c_int, [_c_u32, _c_u32, _vp, _vp, _vp, _c_u32, _c_u32, _c_u32, _vp, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "ubs_isect_offset_encode": (_c_int, [_c_i64, _vp, _c_u32, _c_u32, _c_u32, _vp, _vp]),
    "ubs_rasterize_workspace_bytes": (_c_size, [_c_u32, _c_u32, _c_int]),
    "ubs_rasterize_workspace_bytes_tiles": (_c_size, [_c_u32, _c_u32, _c_u32, _c_u32, _c_int]),
    "ubs_rasterize_to_pixels_fwd": (
        _c_int, [_c_u32, _c_u32, _c_i64, _c_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_u32, _c_u32, _c_u32,
                 _c_u32, _c_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "ubs_rasterize_to_pixels_fwd_prepacked": (
        _c_int, [_c_u32, _c_u32, _c_i64, _c_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_u32, _c_u32, _c_u32,
                 _c_u32, _c_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "ubs_rasterize_to_pixels_fwd_dev": (
        _c_int, [_c_u32, _c_u32, _c_i64, _vp, _c_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_u32, _c_u32, _c_u32,
                 _c_u32, _c_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "ubs_rasterize_to_pixels_bwd": (
        _c_int, [_c_u32, _c_u32, _c_i64, _c_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_u32, _c_u32, _c_u32,
                 _c_u32, _c_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "ubs_transmittance_checkpoint_floats": (_c_size, [_c_u32, _c_i64, _c_u32, _c_u32, _c_u32]),
    "ubs_rasterize_to_pixels_fwd_ckpt": (
        _c_int, [_c_u32, _c_u32, _c_i64, _c_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_u32, _c_u32, _c_u32,
                 _c_u32, _c_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_size, _vp, _c_size, _vp]),
    "ubs_rasterize_to_pixels_bwd_ckpt": (
        _c_int, [_c_u32, _c_u32, _c_i64, _c_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_u32, _c_u32, _c_u32,
                 _c_u32, _c_u32, _vp, _vp, _vp, _vp, _vp, _vp, _c_size, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                 _c_size, _vp]),
    "ubs_rasterize_to_pixels_bwd_reuse": (
        _c_int, [_c_u32, _c_u32, _c_i64, _c_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_u32, _c_u32, _c_u32,
                 _c_u32, _c_u32, _vp, _vp, _vp, _vp, _vp, _vp, _c_size, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                 _c_size, _vp, _c_size, _vp]),
    "ubs_image_loss_workspace_bytes": (_c_size, [_c_i64, _c_i64, _c_i64, _c_i64]),
    "ubs_image_loss_fwd": (_c_int, [_c_i64, _c_i64, _c_i64, _c_i64, _vp, _vp, _c_f32, _vp, _vp, _c_size, _vp]),
    "ubs_image_loss_bwd": (_c_int, [_c_i64, _c_i64, _c_i64, _c_i64, _vp, _vp, _c_f32, _vp, _vp, _c_size, _vp, _vp]),
    "ubs_adam_step": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, _vp]),
    "ubs_xyz_noise": (_c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _vp, _vp, _c_f32, _vp]),
    "ubs_mcmc_relocate": (_c_int, [_c_i64, _c_int, _vp, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _c_f32, _vp]),
    "ubs_set_backward_fp64_transmittance": (_c_int, [_c_int]),
    "ubs_set_cond_backward_exact": (_c_int, [_c_int]),
}


class _Lib:
    """ctypes handle with typed entry points; raises RuntimeError on failure."""

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise ImportError(
                f"libubs_hip.so not found at {path}; build it with "
                "`python universal-beta-splatting_amd/build.py` (gfx950 HIP library required)")
        self._dll = ctypes.CDLL(path)
        self.path = path
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(self._dll, name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def last_error(self) -> str:
        msg = self._dll.ubs_last_error()
        return msg.decode() if msg else ""

    def check(self, rc: int, what: str) -> None:
        if rc != 0:
            raise RuntimeError(f"{what} failed (code {rc}): {self.last_error()}")


_lib: Optional[_Lib] = None


def lib() -> _Lib:
    global _lib
    if _lib is None:
        _lib = _Lib(LIB_PATH)
    return _lib


def ptr(t: Optional[torch.Tensor]):
    """Raw device pointer of a contiguous CUDA (HIP) tensor, or None."""
    if t is None:
        return None
    if not t.is_cuda:
        raise RuntimeError("tensor must be a CUDA (HIP) tensor")
    if not t.is_contiguous():
        raise RuntimeError("tensor must be contiguous")
    return _vp(t.data_ptr())


def stream_of(t: torch.Tensor):
    """Current HIP stream on the tensor's device."""
    return _vp(torch.cuda.current_stream(t.device).cuda_stream)


_F32, _I32, _I64 = torch.float32, torch.int32, torch.int64


def require_dtype(dtype: torch.dtype, *tensors: Optional[torch.Tensor]) -> None:
    """The kernels read raw fp32 / int32 / int64 buffers; any other dtype is an
    error, as the reference's data_ptr<T>() checks make it."""
    for t in tensors:
        if t is not None and t.dtype != dtype:
            raise RuntimeError(f"expected scalar type {dtype} but found {t.dtype}")


def require_cuda(*tensors: Optional[torch.Tensor]) -> None:
    for t in tensors:
        if t is None:
            continue
        if not t.is_cuda:
            raise RuntimeError("gsplat (ubs) ops require tensors on a HIP/CUDA device; "
                               "there is no CPU implementation")
        if not t.is_contiguous():
            raise RuntimeError("tensor must be contiguous")
