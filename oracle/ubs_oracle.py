"""TEST INFRASTRUCTURE ONLY -- numpy front end of the CPU oracle.

Loads ``oracle/libubs_oracle.so`` (built from ``oracle/ubs_oracle.cpp`` by
``oracle/Makefile``) through ctypes.  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg may import this module, and only as
the checker; the product path never does.

Every function takes numpy arrays; float32 inputs run the float
instantiation (bit-exact twin of the HIP kernels for projection, tile
intersection, sort and ranges), float64 inputs run the FP64 oracle.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libubs_oracle.so")
_lib = None

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)


def build(force: bool = False) -> str:
    """Compile the oracle library with make (g++)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle inputs must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def _fn(name: str, dtype) -> ctypes._CFuncPtr:
    suf = "_f64" if np.dtype(dtype) == np.float64 else "_f32"
    f = getattr(lib(), name + suf)
    return f


def _c(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


def num_threads() -> int:
    return int(lib().ora_num_threads())


# ----------------------------------------------------------------------------
# l_triangle_to_rotmat, rot_scale_l_triangle_to_covar
# ----------------------------------------------------------------------------
def rotmat_fwd(l):
    dt = l.dtype
    l = _c(l, dt)
    N = l.shape[0]
    R = np.empty((N, 3, 3), dt)
    _fn("ora_rotmat_fwd", dt)(ctypes.c_int64(N), _ptr(l), _ptr(R))
    return R


def rotmat_bwd(vR):
    dt = vR.dtype
    vR = _c(vR, dt)
    N = vR.shape[0]
    out = np.empty((N, 3), dt)
    _fn("ora_rotmat_bwd", dt)(ctypes.c_int64(N), _ptr(vR), _ptr(out))
    return out


def covar_fwd(rot, scale, ltri, rest_i, rest_j, spatial_block=False):
    dt = scale.dtype
    rot, scale, ltri = _c(rot, dt), _c(scale, dt), _c(ltri, dt)
    ri, rj = _c(rest_i, np.int32), _c(rest_j, np.int32)
    N, D = scale.shape
    dim = 3 if (spatial_block or D == 3) else D
    out = np.empty((N, dim, dim), dt)
    _fn("ora_covar_fwd", dt)(ctypes.c_int64(N), ctypes.c_int64(D), _ptr(rot), _ptr(scale),
                             _ptr(ltri), _ptr(ri), _ptr(rj), ctypes.c_int64(ri.size),
                             ctypes.c_int(int(spatial_block)), _ptr(out))
    return out


def covar_bwd(rot, scale, ltri, rest_i, rest_j, spatial_block, v_covar):
    dt = scale.dtype
    rot, scale, ltri, v_covar = _c(rot, dt), _c(scale, dt), _c(ltri, dt), _c(v_covar, dt)
    ri, rj = _c(rest_i, np.int32), _c(rest_j, np.int32)
    N, D = scale.shape
    vr = np.empty((N, 3, 3), dt)
    vs = np.empty((N, D), dt)
    vl = np.empty((N, D * (D - 1) // 2), dt)
    _fn("ora_covar_bwd", dt)(ctypes.c_int64(N), ctypes.c_int64(D), _ptr(rot), _ptr(scale),
                             _ptr(ltri), _ptr(ri), _ptr(rj), ctypes.c_int64(ri.size),
                             ctypes.c_int(int(spatial_block)), _ptr(v_covar), _ptr(vr), _ptr(vs),
                             _ptr(vl))
    return vr, vs, vl


# ----------------------------------------------------------------------------
# cond_mean_convariance_opacity
# ----------------------------------------------------------------------------
def cond_fwd(means, covars, opacities, betas, query):
    dt = means.dtype
    means, covars, opacities, betas, query = (_c(a, dt) for a in (means, covars, opacities, betas, query))
    N, D = means.shape
    om = np.empty((N, 3), dt)
    oV = np.empty((N, 3, 3), dt)
    oo = np.empty((N, 1), dt)
    _fn("ora_cond_fwd", dt)(ctypes.c_int64(N), ctypes.c_int64(D), _ptr(means), _ptr(covars),
                            _ptr(opacities), _ptr(betas), _ptr(query), _ptr(om), _ptr(oV), _ptr(oo))
    return om, oV, oo


def cond_bwd(means, covars, opacities, betas, query, v_m, v_V, v_o, exact=False):
    """exact=False: the reference's gradient (its Cholesky backward solves with
    L where L^T is needed); exact=True: the true gradient."""
    dt = means.dtype
    arrs = [_c(a, dt) for a in (means, covars, opacities, betas, query, v_m, v_V, v_o)]
    N, D = arrs[0].shape
    C = D - 3
    gm = np.empty((N, D), dt)
    gV = np.empty((N, D, D), dt)
    go = np.empty((N, 1), dt)
    gb = np.empty((N, C), dt)
    _fn("ora_cond_bwd", dt)(ctypes.c_int64(N), ctypes.c_int64(D), *[_ptr(a) for a in arrs],
                            _ptr(gm), _ptr(gV), _ptr(go), _ptr(gb), ctypes.c_int(int(exact)))
    return gm, gV, go, gb


# ----------------------------------------------------------------------------
# fully_fused_projection
# ----------------------------------------------------------------------------
def proj_fwd(means, covars, quats, scales, viewmats, Ks, width, height, eps2d=0.3,
             near_plane=0.01, far_plane=1e10, radius_clip=0.0, calc_compensations=False,
             ortho=False):
    dt = means.dtype
    means, covars, quats, scales, viewmats, Ks = (_c(a, dt) for a in (means, covars, quats, scales, viewmats, Ks))
    C, N = viewmats.shape[0], means.shape[0]
    radii = np.zeros((C, N), np.int32)
    means2d = np.zeros((C, N, 2), dt)
    depths = np.zeros((C, N), dt)
    conics = np.zeros((C, N, 3), dt)
    comp = np.zeros((C, N), dt) if calc_compensations else None
    ct = ctypes.c_double if dt == np.float64 else ctypes.c_float
    _fn("ora_proj_fwd", dt)(ctypes.c_int64(C), ctypes.c_int64(N), _ptr(means), _ptr(covars),
                            _ptr(quats), _ptr(scales), _ptr(viewmats), _ptr(Ks),
                            ctypes.c_int32(width), ctypes.c_int32(height), ct(eps2d),
                            ct(near_plane), ct(far_plane), ct(radius_clip), ctypes.c_int(int(ortho)),
                            _ptr(radii), _ptr(means2d), _ptr(depths), _ptr(conics), _ptr(comp))
    return radii, means2d, depths, conics, comp


def proj_bwd(means, covars, quats, scales, viewmats, Ks, width, height, eps2d, radii, conics,
             comp, v_means2d, v_depths, v_conics, v_comp=None, ortho=False,
             viewmats_requires_grad=False):
    dt = means.dtype
    means, covars, quats, scales, viewmats, Ks, conics, comp, v_means2d, v_depths, v_conics, v_comp = (
        _c(a, dt) for a in (means, covars, quats, scales, viewmats, Ks, conics, comp, v_means2d,
                            v_depths, v_conics, v_comp))
    radii = _c(radii, np.int32)
    C, N = viewmats.shape[0], means.shape[0]
    v_means = np.zeros((N, 3), dt)
    v_covars = np.zeros((N, 6), dt) if covars is not None else None
    v_quats = np.zeros((N, 4), dt) if covars is None else None
    v_scales = np.zeros((N, 3), dt) if covars is None else None
    v_viewmats = np.zeros((C, 4, 4), dt) if viewmats_requires_grad else None
    ct = ctypes.c_double if dt == np.float64 else ctypes.c_float
    _fn("ora_proj_bwd", dt)(ctypes.c_int64(C), ctypes.c_int64(N), _ptr(means), _ptr(covars),
                            _ptr(quats), _ptr(scales), _ptr(viewmats), _ptr(Ks),
                            ctypes.c_int32(width), ctypes.c_int32(height), ct(eps2d),
                            ctypes.c_int(int(ortho)), _ptr(radii), _ptr(conics), _ptr(comp),
                            _ptr(v_means2d), _ptr(v_depths), _ptr(v_conics), _ptr(v_comp),
                            _ptr(v_means), _ptr(v_covars), _ptr(v_quats), _ptr(v_scales),
                            _ptr(v_viewmats))
    return v_means, v_covars, v_quats, v_scales, v_viewmats


# ----------------------------------------------------------------------------
# isect_tiles / sort / isect_offset_encode (float32 structure only)
# ----------------------------------------------------------------------------
def isect_tiles(means2d, radii, depths, tile_size, tile_width, tile_height, sort=True):
    means2d = _c(means2d, np.float32)
    radii = _c(radii, np.int32)
    depths = _c(depths, np.float32)
    C, N = radii.shape
    tpg = np.zeros((C, N), np.int32)
    cum = np.zeros((C * N,), np.int64)
    L = lib()
    L.ora_isect_count.restype = ctypes.c_int64
    n = int(L.ora_isect_count(ctypes.c_int64(C), ctypes.c_int64(N), _ptr(means2d), _ptr(radii),
                              ctypes.c_uint32(tile_size), ctypes.c_uint32(tile_width),
                              ctypes.c_uint32(tile_height), _ptr(tpg), _ptr(cum)))
    ids = np.zeros((n,), np.int64)
    flat = np.zeros((n,), np.int32)
    L.ora_isect_emit(ctypes.c_int64(C), ctypes.c_int64(N), _ptr(means2d), _ptr(radii),
                     _ptr(depths), _ptr(cum), ctypes.c_uint32(tile_size),
                     ctypes.c_uint32(tile_width), ctypes.c_uint32(tile_height),
                     ctypes.c_int(int(sort)), _ptr(ids), _ptr(flat))
    return tpg, ids, flat


def sort_pairs(keys, vals, begin_bit, end_bit):
    keys = np.array(keys, dtype=np.int64, copy=True)
    vals = np.array(vals, dtype=np.int32, copy=True)
    lib().ora_sort_pairs(ctypes.c_int64(keys.size), _ptr(keys), _ptr(vals), ctypes.c_int(begin_bit),
                         ctypes.c_int(end_bit))
    return keys, vals


def isect_offset_encode(isect_ids, n_cameras, tile_width, tile_height):
    isect_ids = _c(isect_ids, np.int64)
    off = np.zeros((n_cameras, tile_height, tile_width), np.int32)
    lib().ora_offsets(ctypes.c_int64(isect_ids.size), _ptr(isect_ids), ctypes.c_int64(n_cameras),
                      ctypes.c_uint32(tile_width), ctypes.c_uint32(tile_height), _ptr(off))
    return off


def tile_bits(tile_width, tile_height):
    return int(tile_width * tile_height).bit_length()


# ----------------------------------------------------------------------------
# rasterize_to_pixels
# ----------------------------------------------------------------------------
def raster_fwd(means2d, conics, colors, opacities, betas, width, height, tile_size,
               isect_offsets, flatten_ids, backgrounds=None, masks=None, margins=False):
    dt = means2d.dtype
    means2d, conics, colors, opacities, betas, backgrounds = (
        _c(a, dt) for a in (means2d, conics, colors, opacities, betas, backgrounds))
    isect_offsets = _c(isect_offsets, np.int32)
    flatten_ids = _c(flatten_ids, np.int32)
    masks = None if masks is None else np.ascontiguousarray(masks, dtype=np.uint8)
    C, N = opacities.shape
    D = colors.shape[-1]
    th, tw = isect_offsets.shape[1:3]
    out_c = np.zeros((C, height, width, D), dt)
    out_a = np.zeros((C, height, width, 1), dt)
    last = np.zeros((C, height, width), np.int32)
    mT = np.zeros((C, height, width), np.float64) if margins else None
    mS = np.zeros((C, height, width), np.float64) if margins else None
    _fn("ora_raster_fwd", dt)(ctypes.c_int64(C), ctypes.c_int64(N),
                              ctypes.c_int64(flatten_ids.size), ctypes.c_int64(D), _ptr(means2d),
                              _ptr(conics), _ptr(colors), _ptr(opacities), _ptr(betas),
                              _ptr(backgrounds), _ptr(masks), ctypes.c_int32(width),
                              ctypes.c_int32(height), ctypes.c_int32(tile_size), ctypes.c_int32(tw),
                              ctypes.c_int32(th), _ptr(isect_offsets), _ptr(flatten_ids),
                              _ptr(out_c), _ptr(out_a), _ptr(last), _ptr(mT), _ptr(mS))
    if margins:
        return out_c, out_a, last, mT, mS
    return out_c, out_a, last


def raster_bwd(means2d, conics, colors, opacities, betas, width, height, tile_size,
               isect_offsets, flatten_ids, render_alphas, last_ids, v_render_colors,
               v_render_alphas, backgrounds=None, masks=None, T_final=None):
    dt = means2d.dtype
    means2d, conics, colors, opacities, betas, backgrounds, render_alphas, v_render_colors, v_render_alphas, T_final = (
        _c(a, dt) for a in (means2d, conics, colors, opacities, betas, backgrounds, render_alphas,
                            v_render_colors, v_render_alphas, T_final))
    isect_offsets = _c(isect_offsets, np.int32)
    flatten_ids = _c(flatten_ids, np.int32)
    last_ids = _c(last_ids, np.int32)
    masks = None if masks is None else np.ascontiguousarray(masks, dtype=np.uint8)
    C, N = opacities.shape
    D = colors.shape[-1]
    th, tw = isect_offsets.shape[1:3]
    v_means2d = np.zeros((C, N, 2), dt)
    v_conics = np.zeros((C, N, 3), dt)
    v_colors = np.zeros((C, N, D), dt)
    v_opac = np.zeros((C, N), dt)
    v_betas = np.zeros((C, N), dt)
    _fn("ora_raster_bwd", dt)(ctypes.c_int64(C), ctypes.c_int64(N),
                              ctypes.c_int64(flatten_ids.size), ctypes.c_int64(D), _ptr(means2d),
                              _ptr(conics), _ptr(colors), _ptr(opacities), _ptr(betas),
                              _ptr(backgrounds), _ptr(masks), ctypes.c_int32(width),
                              ctypes.c_int32(height), ctypes.c_int32(tile_size), ctypes.c_int32(tw),
                              ctypes.c_int32(th), _ptr(isect_offsets), _ptr(flatten_ids),
                              _ptr(render_alphas), _ptr(last_ids), _ptr(v_render_colors),
                              _ptr(v_render_alphas), _ptr(T_final), _ptr(v_means2d),
                              _ptr(v_conics), _ptr(v_colors), _ptr(v_opac), _ptr(v_betas))
    return v_means2d, v_conics, v_colors, v_opac, v_betas


# ----------------------------------------------------------------------------
# rasterization() orchestration (submodules/gsplat/rendering.py:17-232), RGB
# mode, covars given as [N,3,3] (the BetaModel call) or quats/scales.
# ----------------------------------------------------------------------------
TRIU = ([0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2])


def rasterization(means, covars33, opacities, betas, colors, viewmats, Ks, width, height,
                  near_plane=0.01, far_plane=1e10, radius_clip=0.0, eps2d=0.3, tile_size=16,
                  backgrounds=None, structure=None, quats=None, scales=None, tile_mask=None):
    """Forward of rendering.py:rasterization for one dtype.

    ``structure`` lets an FP64 evaluation reuse the float32 pipeline's tile
    lists (isect_offsets, flatten_ids) so both blend in the same order.
    ``tile_mask`` [C,th,tw] blends only the tiles it marks (the blend's
    tile masks, rasterize_to_pixels_fwd.cu:47-56); the others get the
    background.  Returns (render_colors, render_alphas, meta).
    """
    dt = means.dtype
    C = viewmats.shape[0]
    covars6 = None if covars33 is None else np.ascontiguousarray(covars33[:, TRIU[0], TRIU[1]])
    radii, means2d, depths, conics, _ = proj_fwd(means, covars6, quats, scales, viewmats, Ks, width,
                                                 height, eps2d, near_plane, far_plane, radius_clip)
    opac = np.ascontiguousarray(np.broadcast_to(opacities[None], (C,) + opacities.shape))
    bet = np.ascontiguousarray(np.broadcast_to(betas[None], (C,) + betas.shape))
    cols = colors if colors.ndim == 3 else np.ascontiguousarray(np.broadcast_to(colors[None], (C,) + colors.shape))
    tw = math.ceil(width / float(tile_size))
    th = math.ceil(height / float(tile_size))
    if structure is None:
        tpg, ids, flat = isect_tiles(means2d.astype(np.float32), radii, depths.astype(np.float32),
                                     tile_size, tw, th)
        offs = isect_offset_encode(ids, C, tw, th)
    else:
        tpg, ids, flat, offs = structure
    out_c, out_a, last = raster_fwd(means2d, conics, cols, opac, bet, width, height, tile_size, offs,
                                    flat, backgrounds, masks=tile_mask)
    meta = dict(radii=radii, means2d=means2d, depths=depths, conics=conics, opacities=opac,
                betas=bet, colors=cols, tile_width=tw, tile_height=th, tiles_per_gauss=tpg,
                isect_ids=ids, flatten_ids=flat, isect_offsets=offs, last_ids=last,
                covars6=covars6, width=width, height=height, tile_size=tile_size, n_cameras=C)
    return out_c, out_a, meta


# ----------------------------------------------------------------------------
# Full Beta step (BetaModel.render + L1 loss + backward to the leaf
# parameters), scene/beta_model.py:36-52,103-159,660-722 and
# rendering.py:17-232, composed from the functions above.
# ----------------------------------------------------------------------------
def _softplus(x):
    return np.logaddexp(0.0, x)


def _sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def _quat_R(q):
    qn = q / np.linalg.norm(q, axis=-1, keepdims=True)
    w, x, y, z = qn[:, 0], qn[:, 1], qn[:, 2], qn[:, 3]
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                  2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                  2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)
    return R.reshape(-1, 3, 3), qn


def _quat_R_vjp(q, qn, G):
    """d loss / d q for R(q / |q|) given G = d loss / d R."""
    w, x, y, z = qn[:, 0], qn[:, 1], qn[:, 2], qn[:, 3]
    g = G.reshape(-1, 9)
    zero = np.zeros_like(w)
    dw = np.stack([zero, -2 * z, 2 * y, 2 * z, zero, -2 * x, -2 * y, 2 * x, zero], -1)
    dx = np.stack([zero, 2 * y, 2 * z, 2 * y, -4 * x, -2 * w, 2 * z, 2 * w, -4 * x], -1)
    dy = np.stack([-4 * y, 2 * x, 2 * w, 2 * x, zero, 2 * z, -2 * w, 2 * z, -4 * y], -1)
    dz = np.stack([-4 * z, -2 * w, 2 * x, 2 * w, -4 * z, 2 * y, 2 * x, 2 * y, zero], -1)
    gqn = np.stack([(g * dw).sum(-1), (g * dx).sum(-1), (g * dy).sum(-1), (g * dz).sum(-1)], -1)
    nrm = np.linalg.norm(q, axis=-1, keepdims=True)
    return (gqn - (gqn * qn).sum(-1, keepdims=True) * qn) / nrm


def beta_step(leaves: dict, viewmat, K, width, height, target, timestamp=0.0, dtype=np.float64,
              tile_size=16, structure=None, want_grads=True, exact_cond=False, fixed_query=None,
              tile_mask=None):
    """One forward+backward of the Beta rasteriser path on the CPU.

    leaves: xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, rgb,
    quats (numpy).  Loss = mean |image - target|.  Returns (image [H,W,3],
    grads dict, meta).  ``structure`` (from a float32 run's meta) makes an
    FP64 evaluation blend in the float32 tile lists.  ``tile_mask``
    [th,tw] restricts the blend and the loss to the marked tiles (the loss
    is then the mean over their pixels).
    """
    dt = dtype
    L = {k: np.asarray(v, dtype=np.float64) for k, v in leaves.items()}
    D = L["scale_raw"].shape[1]
    scale = _softplus(L["scale_raw"])
    opac = _sigmoid(L["opacity_raw"])
    beta = 4.0 * np.exp(L["beta_raw"])
    rot, qn = _quat_R(L["quats"])
    ri, rj = [], []
    for r in range(D):
        for c in range(r):
            if r >= 3 or c >= 3:
                ri.append(r)
                rj.append(c)
    ri = np.asarray(ri, np.int32)
    rj = np.asarray(rj, np.int32)
    c_ = lambda a: np.ascontiguousarray(a, dtype=dt)
    cov = covar_fwd(c_(rot), c_(scale), c_(L["l_triangle"]), ri, rj)
    mean = np.concatenate([L["xyz"], L["mean_rest"]], 1)
    vm = np.asarray(viewmat, np.float64)
    cc = -vm[:3, :3].T @ vm[:3, 3]
    vd = L["xyz"] - cc
    vd = vd / np.linalg.norm(vd, axis=1, keepdims=True)
    q = vd if D == 6 else np.concatenate([vd, np.full((vd.shape[0], 1), timestamp)], 1)
    if fixed_query is not None:  # (the query carries no gradient; FD checks hold it fixed)
        q = np.asarray(fixed_query, np.float64)
    bcond = np.ascontiguousarray(beta[:, 1:])
    m, V, o = cond_fwd(c_(mean), cov, c_(opac), c_(bcond), c_(q))
    tmask = None if tile_mask is None else np.ascontiguousarray(np.asarray(tile_mask, bool)[None])
    img, alpha, meta = rasterization(m, V, c_(o[:, 0]), c_(beta[:, 0]), c_(L["rgb"]), c_(vm[None]),
                                     c_(np.asarray(K)[None]), width, height, tile_size=tile_size,
                                     structure=structure, tile_mask=tmask)
    meta["query"] = q
    if not want_grads:
        return img[0], None, meta
    H, W_ = height, width
    if tmask is None:
        v_img = np.sign(img[0].astype(np.float64) - target) / (H * W_ * 3)
    else:
        pm = np.kron(tmask[0], np.ones((tile_size, tile_size), bool))[:H, :W_]
        v_img = np.sign(img[0].astype(np.float64) - target) * pm[..., None] / (pm.sum() * 3)
    v_img = c_(v_img[None])
    T_final = c_(1.0 - alpha[..., 0])
    vm2, vcn, vcol, vop, vbe = raster_bwd(meta["means2d"], meta["conics"], meta["colors"], meta["opacities"],
                                          meta["betas"], width, height, tile_size, meta["isect_offsets"],
                                          meta["flatten_ids"], alpha, meta["last_ids"], v_img,
                                          np.zeros_like(alpha), masks=tmask, T_final=T_final)
    v_means, v_cov6, _, _, _ = proj_bwd(m, meta["covars6"], None, None, c_(vm[None]), c_(np.asarray(K)[None]),
                                        width, height, 0.3, meta["radii"], meta["conics"], None, vm2,
                                        np.zeros_like(meta["depths"]), vcn)
    v_V = np.zeros((m.shape[0], 3, 3), dt)
    v_V[:, TRIU[0], TRIU[1]] = v_cov6
    gm, gV, go, gb = cond_bwd(c_(mean), cov, c_(opac), c_(bcond), c_(q), v_means, v_V, c_(vop[0][:, None]),
                              exact=exact_cond)
    v_rot, v_scale, v_lt = covar_bwd(c_(rot), c_(scale), c_(L["l_triangle"]), ri, rj, False, gV)
    grads = dict(
        xyz=gm[:, :3], mean_rest=gm[:, 3:], scale_raw=v_scale * _sigmoid(L["scale_raw"]),
        l_triangle=v_lt, opacity_raw=go * opac * (1.0 - opac),
        beta_raw=np.concatenate([vbe[0][:, None], gb], 1) * beta, rgb=vcol[0],
        quats=_quat_R_vjp(L["quats"], qn, v_rot.astype(np.float64)))
    return img[0], grads, meta


# ---------------------------------------------------------------------------
# Image loss of the training step: (1 - lambda) L1 + lambda (1 - SSIM)
# (train.py:117-121; SSIM as utils/loss_utils.py:26-80)
# ---------------------------------------------------------------------------
def gauss_window_f32(window_size: int = 11, sigma: float = 1.5) -> np.ndarray:
    """loss_utils.py:26-33: float32 weights exp(-(x - 5)^2 / (2 sigma^2)),
    divided by their float32 sum."""
    g = np.array([math.exp(-((x - window_size // 2) ** 2) / float(2 * sigma ** 2)) for x in range(window_size)],
                 dtype=np.float32)
    return g / g.sum(dtype=np.float32)


def _ssim_map_torch(x, y, g):
    """loss_utils.py:56-80 on [B, C, H, W] fp64 tensors: 2-D window outer(g, g)
    correlated per channel with zero padding 5."""
    import torch
    import torch.nn.functional as F

    C = x.shape[1]
    w = torch.outer(g, g)[None, None].expand(C, 1, g.numel(), g.numel()).contiguous()
    pad = g.numel() // 2
    mu1 = F.conv2d(x, w, padding=pad, groups=C)
    mu2 = F.conv2d(y, w, padding=pad, groups=C)
    mu1_sq, mu2_sq, mu1_mu2 = mu1 * mu1, mu2 * mu2, mu1 * mu2
    s1 = F.conv2d(x * x, w, padding=pad, groups=C) - mu1_sq
    s2 = F.conv2d(y * y, w, padding=pad, groups=C) - mu2_sq
    s12 = F.conv2d(x * y, w, padding=pad, groups=C) - mu1_mu2
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    return ((2 * mu1_mu2 + C1) * (2 * s12 + C2)) / ((mu1_sq + mu2_sq + C1) * (s1 + s2 + C2))


def ssim_map(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """SSIM map (fp64) of images [B, H, W, C] (the rasteriser's layout)."""
    import torch

    g = torch.tensor(gauss_window_f32(), dtype=torch.float64)
    tx = torch.tensor(np.asarray(x, np.float64)).permute(0, 3, 1, 2)
    ty = torch.tensor(np.asarray(y, np.float64)).permute(0, 3, 1, 2)
    return _ssim_map_torch(tx, ty, g).permute(0, 2, 3, 1).numpy()


def image_loss(x: np.ndarray, y: np.ndarray, lambda_dssim: float):
    """(loss, d loss / d x) in fp64 for images [B, H, W, C]:
    (1 - lambda) mean|x - y| + lambda (1 - mean SSIM), the mean over every
    element of the batch (train.py:117-128 averages equal-size views); the
    gradient by autograd of this restatement (sign(0) = 0, as torch's abs)."""
    import torch

    g = torch.tensor(gauss_window_f32(), dtype=torch.float64)
    tx = torch.tensor(np.asarray(x, np.float64), requires_grad=True)
    ty = torch.tensor(np.asarray(y, np.float64))
    loss = (1.0 - lambda_dssim) * (tx - ty).abs().mean()
    if lambda_dssim > 0:
        s = _ssim_map_torch(tx.permute(0, 3, 1, 2), ty.permute(0, 3, 1, 2), g)
        loss = loss + lambda_dssim * (1.0 - s.mean())
    loss.backward()
    return float(loss.detach()), tx.grad.numpy()


# ---------------------------------------------------------------------------
# Step after the backward: Adam (scene/beta_model.py:239-276 builds
# torch.optim.Adam(groups, lr=0.0, eps=1e-15)), covariance-shaped position
# noise (train.py:159-168) and the MCMC row relocation
# (beta_model.py:553-658).
# ---------------------------------------------------------------------------
def adam_step(p, g, m, v, lr, step, beta1=0.9, beta2=0.999, eps=1e-15):
    """One Adam update of a tensor in fp64 (torch.optim.Adam, no weight
    decay, no amsgrad): returns (p', m', v') for the step count ``step``
    (>= 1, the count after this update)."""
    p, g, m, v = (np.asarray(a, np.float64) for a in (p, g, m, v))
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    bc1 = 1.0 - beta1 ** step
    bc2 = 1.0 - beta2 ** step
    p = p - (lr / bc1) * m / (np.sqrt(v) / math.sqrt(bc2) + eps)
    return p, m, v


def xyz_covariance(l_triangle, scale_raw):
    """get_xyz_covariance (beta_model.py:143-152): the spatial 3x3 block of
    the covariance, through the oracle's rotmat_fwd / covar_fwd."""
    lt = np.asarray(l_triangle, np.float64)
    s = _softplus(np.asarray(scale_raw, np.float64))
    D = s.shape[1]
    R = rotmat_fwd(np.ascontiguousarray(lt[:, :3]))
    ti, tj = np.tril_indices(D, -1)
    keep = (ti >= 3) | (tj >= 3)
    return covar_fwd(R, s, lt, ti[keep].astype(np.int32), tj[keep].astype(np.int32), spatial_block=True)


def xyz_noise(xyz, l_triangle, scale_raw, opacity_raw, normal, noise_lr, xyz_lr):
    """train.py:159-168 in fp64: xyz + Sigma_xyz (z (1 - o)^100 noise_lr xyz_lr)."""
    o = _sigmoid(np.asarray(opacity_raw, np.float64))
    z = np.asarray(normal, np.float64) * (1.0 - o) ** 100 * noise_lr * xyz_lr
    S = xyz_covariance(l_triangle, scale_raw)
    return np.asarray(xyz, np.float64) + np.einsum("nij,nj->ni", S, z)


def split_opacity_raw(opacity_raw, ratio, min_opacity=0.005):
    """_update_params (beta_model.py:552-570): the pre-activation opacity of
    each of ratio + 1 copies, logit(clamp(1 - (1 - o)^(1/(ratio+1))))."""
    o = _sigmoid(np.asarray(opacity_raw, np.float64))
    on = 1.0 - (1.0 - o) ** (1.0 / (np.asarray(ratio, np.float64) + 1.0))
    on = np.clip(on, min_opacity, 1.0 - float(np.finfo(np.float32).eps))
    return np.log(on / (1.0 - on))


def mcmc_relocate(leaves: dict, exp_avg: dict, exp_avg_sq: dict, src, dst, ratio, opacity_key="opacity_raw"):
    """The row updates of relocate_gs / add_new_gs (beta_model.py:580-658)
    for sampled (src, dst, ratio): dst rows take src rows (opacity split),
    then src rows take the split opacity and zeroed moments.  Arrays are
    copied, fp64."""
    L = {k: np.array(v, np.float64) for k, v in leaves.items()}
    M = {k: np.array(v, np.float64) for k, v in exp_avg.items()}
    V = {k: np.array(v, np.float64) for k, v in exp_avg_sq.items()}
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    new_o = split_opacity_raw(L[opacity_key][src, 0], ratio)
    for k in L:
        L[k][dst] = L[k][src]
    L[opacity_key][dst, 0] = new_o
    L[opacity_key][src, 0] = new_o
    for k in M:
        M[k][src] = 0.0
        V[k][src] = 0.0
    return L, M, V
