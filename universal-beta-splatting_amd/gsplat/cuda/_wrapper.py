"""Python API of the Beta-splat rasteriser ops, backed by libubs_hip.so.

Same function names, arguments, shapes, autograd behaviour and error
behaviour as the reference's submodules/gsplat/cuda/_wrapper.py (the entry
points BetaModel and rasterization() call).  Each op allocates its outputs
and workspaces through torch's caching allocator and hands raw pointers and
the current HIP stream to the C ABI (include/ubs_capi.h).
"""
from __future__ import annotations

import math
from typing import Optional, Tuple

import torch
from torch import Tensor

from ._backend import _F32, _I32, _I64, lib, ptr, require_cuda, require_dtype, stream_of


class LaunchTimer:
    """Records a pair of HIP events (torch.cuda.Event, current stream -- the
    stream every op launches on) around each C-ABI call whose name is in
    ``names`` (all when None).  Install with ``set_launch_timer``."""

    def __init__(self, names=None):
        self.names = None if names is None else set(names)
        self.events = []  # (name, start, end)

    def wants(self, name: str) -> bool:
        return self.names is None or name in self.names

    def durations_ms(self):
        torch.cuda.synchronize()
        out = {}
        for name, s, e in self.events:
            out.setdefault(name, []).append(s.elapsed_time(e))
        return out


_timer: Optional[LaunchTimer] = None


def set_launch_timer(timer: Optional[LaunchTimer]) -> None:
    global _timer
    _timer = timer


def _call(name: str, *args) -> None:
    L = lib()
    t = _timer
    if t is not None and t.wants(name):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        rc = getattr(L, name)(*args)
        e.record()
        t.events.append((name, s, e))
    else:
        rc = getattr(L, name)(*args)
    L.check(rc, name)


# ----------------------------------------------------------------------------
# Public API (reference: _wrapper.py:18-55, 185-491)
# ----------------------------------------------------------------------------
def cond_mean_convariance_opacity(
    means: Tensor,  # [N, D]
    covars: Tensor,  # [N, D, D]
    opacities: Tensor,  # [N, 1]
    betas: Tensor,  # [N, C]
    query: Tensor,  # [N, C]
) -> Tuple[Tensor, Tensor, Tensor]:
    """Condition each D-dim primitive on its C = D-3 query dims.

    Returns the conditioned mean [N,3], covariance [N,3,3] and opacity [N,1].
    reference: _wrapper.py:18-31; the query gets no gradient (:611).
    """
    return _CondMeanConvarianceOpacity.apply(
        means.contiguous(), covars.contiguous(), opacities.contiguous(), betas.contiguous(),
        query.contiguous())


def l_triangle_to_rotmat(l_triangle: Tensor) -> Tensor:
    """First-order rotation R = I + A from the skew entries [N,3] (reference: _wrapper.py:34-36)."""
    assert l_triangle.shape[1] == 3, l_triangle.shape
    return _LTriangleToRotmat.apply(l_triangle.contiguous())


def beta_activations(
    xyz: Tensor,  # [N, 3]
    mean_rest: Tensor,  # [N, D-3]
    scale_raw: Tensor,  # [N, D]
    opacity_raw: Tensor,  # [N, 1]
    beta_raw: Tensor,  # [N, D-2]
    quats: Optional[Tensor],  # [N, 4] or None
    cam_pos: Optional[Tensor],  # [3] or None (no query)
    timestamp: float = 0.0,
) -> Tuple[Tensor, Tensor, Tensor, Tensor, Optional[Tensor], Optional[Tensor]]:
    """BetaModel's activations, view/time query and quaternion rotation in one
    kernel: (mean [N,D], scale [N,D], opacity [N,1], beta [N,D-2],
    query [N,D-3] or None, rot [N,3,3] or None).  The query has no gradient.
    reference: scene/beta_model.py:36-52,103-121 (activations), :675-690 (query)"""
    D = scale_raw.shape[1]
    if cam_pos is not None and D not in (6, 7):
        raise NotImplementedError("Only implemented for 6D or 7D query")
    return _BetaActivations.apply(xyz.contiguous(), mean_rest.contiguous(), scale_raw.contiguous(),
                                  opacity_raw.contiguous(), beta_raw.contiguous(),
                                  None if quats is None else quats.contiguous(),
                                  None if cam_pos is None else cam_pos.contiguous(), float(timestamp))


def rot_scale_l_triangle_to_covar(
    rot: Tensor,  # [N, 3, 3]
    scale: Tensor,  # [N, D]
    l_triangle: Tensor,  # [N, D*(D-1)//2]
    rest_i: Tensor,
    rest_j: Tensor,
    spatial_block: bool = False,
) -> Tensor:
    """Sigma = L L^T (reference: _wrapper.py:39-55)."""
    return _RotScaleLTriangleToCovar.apply(
        rot.contiguous(), scale.contiguous(), l_triangle.contiguous(), rest_i, rest_j, spatial_block)


@torch.no_grad()
def camera_centers(viewmats: Tensor) -> Tensor:
    """World positions [C,3] of world->camera [C,4,4] matrices, -(R^T t) per
    view (the camera centre BetaModel's query is taken from), one launch; the
    same bits as the fused projection's own centre when it is given none."""
    viewmats = viewmats.detach().to(torch.float32).contiguous()
    require_cuda(viewmats)
    C = viewmats.size(0)
    assert viewmats.shape == (C, 4, 4), viewmats.shape
    out = torch.empty((C, 3), dtype=torch.float32, device=viewmats.device)
    _call("ubs_camera_centers", C, ptr(viewmats), ptr(out), stream_of(viewmats))
    return out


def beta_project(
    xyz: Tensor,  # [N, 3]
    mean_rest: Tensor,  # [N, D-3]
    scale_raw: Tensor,  # [N, D]
    l_triangle: Tensor,  # [N, D(D-1)/2]
    opacity_raw: Tensor,  # [N, 1]
    beta_raw: Tensor,  # [N, D-2]
    quats: Optional[Tensor],  # [N, 4] or None
    viewmats: Tensor,  # [C, 4, 4]
    Ks: Tensor,  # [C, 3, 3]
    width: int,
    height: int,
    cam_pos: Optional[Tensor] = None,  # [C, 3]; camera_centers(viewmats), taken in the kernel, when None
    timestamps: Optional[Tensor] = None,  # [C] (7-D); 0 when None
    image_sizes: Optional[Tensor] = None,  # int32 [C, 2] (W, H) per view, else width x height
    eps2d: float = 0.3,
    near_plane: float = 0.01,
    far_plane: float = 1e10,
    radius_clip: float = 0.0,
    ortho: bool = False,
    rgb: Optional[Tensor] = None,  # [N, 3], with records
    records: Optional[Tensor] = None,  # a prepacked_blend_workspace: the blend's records are written too
) -> Tuple[Tensor, Tensor, Tensor, Tensor, Tensor, Tensor]:
    """Fused Beta projection of a 6-D or 7-D scene for C views, from the leaf
    parameters (pre-activation, BetaModel's layout) to the per-view 2-D
    primitives: (radii int32 [C,N], means2d [C,N,2], depths [C,N],
    conics [C,N,3], opacities [C,N], betas [C,N]).

    One launch does, per primitive, what BetaModel.render runs per view before
    the blend (scene/beta_model.py:660-722): activations, query, rotation
    (quaternion, or I + skew(l_triangle[:, :3]) when quats is None),
    Sigma = L L^T, conditioning on the view's query, and the EWA projection of
    rasterization() (rendering.py:54-101, covariances read as their upper
    triangle).  With one view every output equals that chain's bit for bit;
    gradients flow to every leaf, summed over the views."""
    D = scale_raw.shape[1]
    if D not in (6, 7):
        raise NotImplementedError("Only implemented for 6D or 7D query")
    return _BetaProject.apply(
        xyz.contiguous(), mean_rest.contiguous(), scale_raw.contiguous(), l_triangle.contiguous(),
        opacity_raw.contiguous(), beta_raw.contiguous(), None if quats is None else quats.contiguous(),
        viewmats.contiguous(), Ks.contiguous(), None if cam_pos is None else cam_pos.detach().contiguous(),
        None if timestamps is None else timestamps.detach().float().contiguous(),
        None if image_sizes is None else image_sizes.to(torch.int32).contiguous(), int(width), int(height),
        float(eps2d), float(near_plane), float(far_plane), float(radius_clip), bool(ortho),
        None if records is None else rgb.detach().contiguous(), records)


def fully_fused_projection(
    means: Tensor,  # [N, 3]
    covars: Optional[Tensor],  # [N, 6] or None
    quats: Optional[Tensor],  # [N, 4] or None
    scales: Optional[Tensor],  # [N, 3] or None
    viewmats: Tensor,  # [C, 4, 4]
    Ks: Tensor,  # [C, 3, 3]
    width: int,
    height: int,
    eps2d: float = 0.3,
    near_plane: float = 0.01,
    far_plane: float = 1e10,
    radius_clip: float = 0.0,
    sparse_grad: bool = False,
    calc_compensations: bool = False,
    ortho: bool = False,
) -> Tuple[Tensor, Tensor, Tensor, Tensor, Tensor]:
    """EWA projection to 2D: (radii [C,N] int32, means2d [C,N,2], depths [C,N],
    conics [C,N,3], compensations [C,N] or None).  reference: _wrapper.py:185-282"""
    C = viewmats.size(0)
    N = means.size(0)
    assert means.size() == (N, 3), means.size()
    assert viewmats.size() == (C, 4, 4), viewmats.size()
    assert Ks.size() == (C, 3, 3), Ks.size()
    means = means.contiguous()
    if covars is not None:
        assert covars.size() == (N, 6), covars.size()
        covars = covars.contiguous()
    else:
        assert quats is not None, "covars or quats is required"
        assert scales is not None, "covars or scales is required"
        assert quats.size() == (N, 4), quats.size()
        assert scales.size() == (N, 3), scales.size()
        quats = quats.contiguous()
        scales = scales.contiguous()
    viewmats = viewmats.contiguous()
    Ks = Ks.contiguous()
    return _FullyFusedProjection.apply(
        means, covars, quats, scales, viewmats, Ks, width, height, eps2d, near_plane, far_plane,
        radius_clip, calc_compensations, ortho, 0)


def fully_fused_projection_covars33(
    means: Tensor,  # [N, 3]
    covars: Tensor,  # [N, 3, 3], upper triangle read
    viewmats: Tensor,
    Ks: Tensor,
    width: int,
    height: int,
    eps2d: float = 0.3,
    near_plane: float = 0.01,
    far_plane: float = 1e10,
    radius_clip: float = 0.0,
    calc_compensations: bool = False,
) -> Tuple[Tensor, Tensor, Tensor, Tensor, Tensor]:
    """fully_fused_projection on the upper triangle of [N,3,3] covariances.

    Same results as gathering covars[..., [0,0,0,1,1,2], [0,1,2,1,2,2]] and
    projecting (rendering.py:54-56,81-97), with the gather fused into the
    kernel; the gradient w.r.t. covars is upper-triangle-only, as the gather's.
    """
    N = means.size(0)
    assert covars.shape == (N, 3, 3), covars.shape
    return _FullyFusedProjection.apply(
        means.contiguous(), covars.contiguous(), None, None, viewmats.contiguous(), Ks.contiguous(), width,
        height, eps2d, near_plane, far_plane, radius_clip, calc_compensations, False, 1)


@torch.no_grad()
def isect_tiles(
    means2d: Tensor,  # [C, N, 2]
    radii: Tensor,  # [C, N]
    depths: Tensor,  # [C, N]
    tile_size: int,
    tile_width: int,
    tile_height: int,
    sort: bool = True,
    n_cameras: Optional[int] = None,
    camera_ids: Optional[Tensor] = None,
    primitive_ids: Optional[Tensor] = None,
) -> Tuple[Tensor, Tensor, Tensor]:
    """Tiles per primitive [C,N] int32, isect ids int64 [n] (cam | tile |
    depth bits) sorted stably on their low 32 + tile_bits + cam_bits bits, and
    flatten ids int32 [n].  reference: _wrapper.py:285-343, isect_tiles.cu:99-285"""
    return isect_tiles_offsets(means2d, radii, depths, tile_size, tile_width, tile_height, sort)[:3]


_SORT_MODES = ("depth_first", "segmented", "onesweep")
_sort_mode = "depth_first"


def set_isect_sort_mode(mode: str) -> None:
    """How isect ids reach their sorted order; every mode gives the same
    ids and values, bit for bit:
      depth_first (default): the (camera, primitive) pairs sorted by depth
        once, entries emitted in that order, then two onesweep passes over
        the camera/tile bits;
      segmented: entries in primitive order, onesweep over the camera/tile
        bits, then a per-tile depth sort in LDS;
      onesweep: onesweep passes over every byte of the ids."""
    global _sort_mode
    if mode not in _SORT_MODES:
        raise ValueError(f"sort mode must be one of {_SORT_MODES}")
    _sort_mode = mode


def isect_sort_mode() -> str:
    return _sort_mode


def set_segmented_depth_sort(enable: bool) -> None:
    """segmented (True) or onesweep (False) sort mode."""
    set_isect_sort_mode("segmented" if enable else "onesweep")


def isect_tiles_offsets(means2d: Tensor, radii: Tensor, depths: Tensor, tile_size: int, tile_width: int,
                        tile_height: int, sort: bool = True):
    """isect_tiles, plus the tile offsets of the sorted ids when the sort
    computed them on the way (else None): (tiles_per_gauss, isect_ids,
    flatten_ids, isect_offsets or None)."""
    state = isect_begin(means2d, radii, depths, tile_size, tile_width, tile_height, sort)
    (n_isects, packed), = isect_read_info([state])  # host sync, as isect_tiles.cu:181
    return isect_finish(state, n_isects, packed)


# Per device: the status of the isect-id sorts that ran after the last host
# read (nonzero = a look-back wait gave up and that sort's order is invalid;
# the sorts set it, ubs_radix_sort_pairs_flag).  The next isect_begin's
# count kernels copy it into its info, so the host read that already sizes
# the next step reports it.
_late_sort_status = {}


def _late_status(dev) -> Tensor:
    t = _late_sort_status.get(dev)
    if t is None:
        t = _late_sort_status[dev] = torch.zeros((1,), dtype=torch.int64, device=dev)
    return t


def isect_read_info(states) -> list:
    """One host read of the info of several isect_begin states: a list of
    (n_isects, packed depth range).  Raises RuntimeError when a radix sort
    since the previous read gave up on a look-back wait (its output order,
    and what was rendered from it, is invalid)."""
    if not states:
        return []
    if all(st.info_host is not None for st in states):
        for st in states:
            st.ready.synchronize()
        infos = torch.stack([st.info_host for st in states])
    else:
        infos = (torch.stack([st.info for st in states]) if len(states) > 1 else states[0].info[None]).cpu()
    if bool((infos[:, 2] != 0).any()):
        for st in states:
            _late_status(st.info.device).zero_()
        raise RuntimeError("radix sort look-back exceeded its spin limit: the isect order of this or the "
                           "previous step is invalid")
    return [(int(r[0]), int(r[1])) for r in infos]


class IsectState:
    """The intersection stage up to the host read of n_isects (see isect_begin)."""
    __slots__ = ("means2d", "radii", "depths", "tile_size", "tile_width", "tile_height", "sort", "tpg", "cum",
                 "info", "info_host", "ready", "order", "prec", "C", "N")


def isect_begin(means2d: Tensor, radii: Tensor, depths: Tensor, tile_size: int, tile_width: int, tile_height: int,
                sort: bool = True) -> IsectState:
    """First half of isect_tiles: tile counts, their prefix, n_isects and the
    depth range into ``state.info`` ([n_isects, packed range] int64, on the
    device) and, for the depth-first sort, the depth order of the pairs.
    Nothing waits for the host; several states can be read in one sync
    before their isect_finish calls.  With the depth-first sort the info is
    copied to pinned host memory before the depth sort is queued, so the
    host read waits for the counts only and the depth sort runs while the
    host queues the emission; that sort's status then joins the status
    word the next host read reports."""
    C, N, _ = means2d.shape
    assert means2d.shape == (C, N, 2), means2d.size()
    assert radii.shape == (C, N), radii.size()
    assert depths.shape == (C, N), depths.size()
    means2d, radii, depths = means2d.contiguous(), radii.contiguous(), depths.contiguous()
    require_cuda(means2d, radii, depths)
    require_dtype(_F32, means2d, depths)
    require_dtype(_I32, radii)
    dev = means2d.device
    L = lib()
    st = stream_of(means2d)
    S = IsectState()
    S.means2d, S.radii, S.depths, S.C, S.N = means2d, radii, depths, C, N
    S.tile_size, S.tile_width, S.tile_height, S.sort = tile_size, tile_width, tile_height, sort
    S.tpg = torch.empty((C, N), dtype=torch.int32, device=dev)
    S.cum = torch.empty((C * N,), dtype=torch.int64, device=dev)
    # [n_isects, (~min depth bits, max depth bits) as two uint32, sort status],
    # written by the count kernels and read in one sync
    S.info = torch.empty((3,), dtype=torch.int64, device=dev)
    ws_bytes = int(L.ubs_isect_count_workspace_bytes(C, N))
    ws = torch.empty((ws_bytes,), dtype=torch.uint8, device=dev)
    depth_first = sort and _sort_mode == "depth_first" and C * N > 0
    # the depth-first sort's input (depth bits, flat ids), written by the count pass
    dkeys = torch.empty((C * N,), dtype=torch.int64, device=dev) if depth_first else None
    dvals = torch.empty((C * N,), dtype=torch.int32, device=dev) if depth_first else None
    # the emission's per-pair reads (centre, radius, depth) in one 16-byte record
    S.prec = torch.empty((C * N, 4), dtype=torch.float32, device=dev) if depth_first else None
    _call("ubs_isect_tiles_count_info", C, N, ptr(means2d), ptr(radii), ptr(depths), tile_size, tile_width,
          tile_height, ptr(S.tpg), ptr(S.cum), ptr(S.info), ptr(_late_status(dev)), ptr(dkeys), ptr(dvals),
          ptr(S.prec), ptr(ws), ws_bytes, st)
    S.order = S.info_host = S.ready = None
    if depth_first:
        # the info leaves for the host first; the depth sort and its scan
        # run during the round trip
        S.info_host = torch.empty((3,), dtype=torch.int64, pin_memory=True)
        S.info_host.copy_(S.info, non_blocking=True)
        S.ready = torch.cuda.Event()
        S.ready.record(torch.cuda.current_stream(dev))
        _, S.order = radix_sort_pairs(dkeys, dvals, 0, 32, status_out=_late_status(dev))
        S.cum = torch.empty((C * N,), dtype=torch.int64, device=dev)
        ows = int(L.ubs_isect_ordered_workspace_bytes(C, N))
        ow = torch.empty((ows,), dtype=torch.uint8, device=dev)
        _call("ubs_isect_scan_ordered", C, N, ptr(S.tpg), ptr(S.order), ptr(S.cum), ptr(ow), ows, st)
    return S


def isect_finish(S: IsectState, n_isects: int, packed: int):
    """Second half of isect_tiles, given the state's host-read info:
    (tiles_per_gauss, isect_ids, flatten_ids, isect_offsets or None)."""
    L = lib()
    C, dev, st = S.C, S.means2d.device, stream_of(S.means2d)
    tw, th, ts = S.tile_width, S.tile_height, S.tile_size
    isect_ids = torch.empty((n_isects,), dtype=torch.int64, device=dev)
    flatten_ids = torch.empty((n_isects,), dtype=torch.int32, device=dev)
    if n_isects and S.order is not None:
        # capped at n_isects: the depth order may come from a sort whose
        # status the host has not read yet
        _call("ubs_isect_tiles_emit_ordered_rec", C, S.N, ptr(S.prec), ptr(S.order), ptr(S.cum), ts, tw, th,
              ptr(isect_ids), ptr(flatten_ids), n_isects, st)
    elif n_isects:
        _call("ubs_isect_tiles_emit", C, S.N, ptr(S.means2d), ptr(S.radii), ptr(S.depths), ptr(S.cum), ts, tw, th,
              ptr(isect_ids), ptr(flatten_ids), st)
    offsets = None
    if n_isects and S.sort:
        end_bit = 32 + int(L.ubs_isect_prefix_bits(C, tw, th))
        late = _late_status(dev)
        if S.order is not None:  # entries already in depth order: stable on the camera/tile bits finishes it
            isect_ids, flatten_ids = radix_sort_pairs(isect_ids, flatten_ids, 32, end_bit, status_out=late)
        elif _sort_mode == "segmented":
            # stable on the camera/tile bits, then per tile on the depth bits
            ids1, flat1 = radix_sort_pairs(isect_ids, flatten_ids, 32, end_bit, status_out=late)
            offsets = isect_offset_encode(ids1, C, tw, th)
            isect_ids, flatten_ids = torch.empty_like(ids1), torch.empty_like(flat1)
            _call("ubs_sort_tile_segments", n_isects, ptr(ids1), ptr(flat1), ptr(isect_ids), ptr(flatten_ids),
                  ptr(offsets), C, tw, th, constant_depth_passes(packed), st)
        else:
            isect_ids, flatten_ids = radix_sort_pairs(isect_ids, flatten_ids, 0, end_bit,
                                                      skip_passes=constant_depth_passes(packed), status_out=late)
    return S.tpg, isect_ids, flatten_ids, offsets


def constant_depth_passes(packed_range: int) -> int:
    """Bit mask of the 8-bit sort passes over depth bits (passes 0-3 of the
    isect-id sort) on which every visible depth agrees, from the packed
    (~min, max) uint32 pair ubs_isect_tiles_count_range writes: byte p is the same
    for every depth in [min, max] when min and max agree on bits >= 8p."""
    lo = ~packed_range & 0xFFFFFFFF
    hi = (packed_range >> 32) & 0xFFFFFFFF
    mask = 0
    if lo <= hi:
        for p in range(4):
            if (lo >> (8 * p)) == (hi >> (8 * p)):
                mask |= 1 << p
    return mask


def _vp_at(t: Tensor, index: int):
    """Raw pointer to element `index` of a contiguous tensor."""
    import ctypes

    return ctypes.c_void_p(t.data_ptr() + index * t.element_size())


def radix_sort_pairs(keys: Tensor, values: Tensor, begin_bit: int, end_bit: int,
                     check_status: bool = False, skip_passes: int = 0,
                     status_out: Optional[Tensor] = None) -> Tuple[Tensor, Tensor]:
    """Stable onesweep radix sort of (int64 key, int32 value) on key bits
    [begin_bit, end_bit); returns new sorted tensors (double-buffered; the
    inputs may be consumed as scratch).  skip_passes: bit p set leaves out
    the 8-bit pass over bits [begin_bit + 8p, +8), which the caller knows
    every key shares (the pass would be the identity).  check_status reads
    the sort's status word now (a host sync); status_out (int64 [1] on the
    device) instead is set to 1 if the sort gives up (never cleared), without
    a sync, for a later read."""
    require_cuda(keys, values)
    assert keys.dtype == torch.int64 and values.dtype == torch.int32
    n = keys.numel()
    L = lib()
    st = stream_of(keys)
    keys_alt = torch.empty_like(keys)
    vals_alt = torch.empty_like(values)
    ws_bytes = int(L.ubs_radix_sort_workspace_bytes(n, begin_bit, end_bit))
    ws = torch.empty((ws_bytes,), dtype=torch.uint8, device=keys.device)
    import ctypes

    sel = ctypes.c_int(0)
    if status_out is not None:
        # a give-up sets the low word of status_out (int64, little-endian) to 1
        assert status_out.dtype == torch.int64 and status_out.is_contiguous()
        _call("ubs_radix_sort_pairs_flag", n, ptr(keys), ptr(keys_alt), ptr(values), ptr(vals_alt), begin_bit,
              end_bit, int(skip_passes), ptr(ws), ws_bytes, ctypes.byref(sel), ptr(status_out), st)
    elif skip_passes:
        _call("ubs_radix_sort_pairs_skip", n, ptr(keys), ptr(keys_alt), ptr(values), ptr(vals_alt), begin_bit,
              end_bit, int(skip_passes), ptr(ws), ws_bytes, ctypes.byref(sel), st)
    else:
        _call("ubs_radix_sort_pairs", n, ptr(keys), ptr(keys_alt), ptr(values), ptr(vals_alt), begin_bit,
              end_bit, ptr(ws), ws_bytes, ctypes.byref(sel), st)
    if check_status:
        status = ctypes.c_int(0)
        L.check(L.ubs_radix_sort_status(ptr(ws), ctypes.byref(status), st), "radix_sort_status")
        if status.value != 0:
            raise RuntimeError("radix sort look-back exceeded its spin limit; result invalid")
    if sel.value:
        return keys_alt, vals_alt
    return keys, values


@torch.no_grad()
def isect_offset_encode(isect_ids: Tensor, n_cameras: int, tile_width: int, tile_height: int) -> Tensor:
    """Per-(camera, tile) start index [C, th, tw] int32.  reference: _wrapper.py:346-363"""
    isect_ids = isect_ids.contiguous()
    require_cuda(isect_ids)
    require_dtype(_I64, isect_ids)
    offsets = torch.empty((n_cameras, tile_height, tile_width), dtype=torch.int32, device=isect_ids.device)
    _call("ubs_isect_offset_encode", isect_ids.numel(), ptr(isect_ids), n_cameras, tile_width, tile_height,
          ptr(offsets), stream_of(isect_ids))
    return offsets


def rasterize_to_pixels(
    means2d: Tensor,  # [C, N, 2]
    conics: Tensor,  # [C, N, 3]
    colors: Tensor,  # [C, N, channels]
    opacities: Tensor,  # [C, N]
    betas: Tensor,  # [C, N]
    image_width: int,
    image_height: int,
    tile_size: int,
    isect_offsets: Tensor,  # [C, tile_height, tile_width]
    flatten_ids: Tensor,  # [n_isects]
    backgrounds: Optional[Tensor] = None,  # [C, channels]
    masks: Optional[Tensor] = None,  # [C, tile_height, tile_width]
    prepacked_workspace: Optional[Tensor] = None,
) -> Tuple[Tensor, Tensor]:
    """Front-to-back Beta blending: (render_colors [C,H,W,channels],
    render_alphas [C,H,W,1]).  reference: _wrapper.py:366-491
    prepacked_workspace: a prepacked_blend_workspace whose records the fused
    projection wrote (beta_project(records=...)) for these inputs."""
    C = isect_offsets.size(0)
    N = means2d.size(1)
    assert means2d.shape == (C, N, 2), means2d.shape
    assert conics.shape == (C, N, 3), conics.shape
    assert colors.shape[:2] == (C, N), colors.shape
    assert opacities.shape == (C, N), opacities.shape
    assert betas.shape == (C, N), betas.shape
    if backgrounds is not None:
        assert backgrounds.shape == (C, colors.shape[-1]), backgrounds.shape
        backgrounds = backgrounds.contiguous()
    if masks is not None:
        assert masks.shape == isect_offsets.shape, masks.shape
        masks = masks.contiguous()
    channels = colors.shape[-1]
    if channels > 513 or channels == 0:
        raise ValueError(f"Unsupported number of color channels: {channels}")
    tile_height, tile_width = isect_offsets.shape[1:3]
    assert tile_height * tile_size >= image_height, \
        f"Assert Failed: {tile_height} * {tile_size} >= {image_height}"
    assert tile_width * tile_size >= image_width, \
        f"Assert Failed: {tile_width} * {tile_size} >= {image_width}"
    render_colors, render_alphas = _RasterizeToPixels.apply(
        means2d.contiguous(), conics.contiguous(), colors.contiguous(), opacities.contiguous(),
        betas.contiguous(), backgrounds, masks, image_width, image_height, tile_size,
        isect_offsets.contiguous(), flatten_ids.contiguous(), prepacked_workspace)
    return render_colors, render_alphas


def set_backward_fp64_transmittance(enable: bool) -> None:
    """Recover transmittance in fp64 in the blend backward (default fp32)."""
    lib().ubs_set_backward_fp64_transmittance(1 if enable else 0)


_t_checkpoint = False


def set_backward_transmittance_checkpoint(enable: bool) -> None:
    """Per-pixel transmittance in the blend backward: False (default)
    recomputes it by division from the final T, as the reference does
    (rasterize_to_pixels_bwd.cu:202-203); True stores it in the forward after
    every batch of tile_size^2 entries (4 B per pixel per batch) and restarts
    each batch of the backward from the stored value."""
    global _t_checkpoint
    _t_checkpoint = bool(enable)


def set_cond_backward_exact(enable: bool) -> None:
    """Conditioning backward: False (default) reproduces the reference's
    gradient, whose Cholesky backward solves with L where L^T is needed
    (cond_mean_convariance_opacity_bwd.cu:508-518); True gives the exact
    gradient of the conditioned opacity w.r.t. the query covariance block."""
    lib().ubs_set_cond_backward_exact(1 if enable else 0)


# ----------------------------------------------------------------------------
# Raw (non-autograd) launchers, used by the autograd classes and benches
# ----------------------------------------------------------------------------
def _mask_ptr(masks: Optional[Tensor]):
    """bool/uint8 tile mask -> pointer to one byte per tile (nonzero = render)."""
    if masks is None:
        return None
    if masks.dtype not in (torch.bool, torch.uint8):
        raise TypeError("masks must be a bool or uint8 tensor")
    return ptr(masks), masks


_packed_records = True


def set_blend_packed_records(enable: bool) -> None:
    """Gather 32-byte packed primitive records in the blend kernels (default)
    or the four separate input arrays (same results, more gather traffic)."""
    global _packed_records
    _packed_records = bool(enable)


_tile_order = True


def set_blend_tile_order(enable: bool) -> None:
    """Deal the blend's tile blocks to the GPU heaviest first (default: the
    workspace gets room for the order, ubs_rasterize_workspace_bytes_tiles) or
    in raster order.  The forward's results are the same either way; the
    backward's float atomics add in another order (same within rounding)."""
    global _tile_order
    _tile_order = bool(enable)


def blend_workspace_bytes(C: int, N: int, tile_width: int, tile_height: int, backward: bool) -> int:
    """Bytes of the blend workspace the wrappers pass (records, and the
    block order unless set_blend_tile_order(False))."""
    L = lib()
    if _tile_order:
        return int(L.ubs_rasterize_workspace_bytes_tiles(C, N, tile_width, tile_height, 1 if backward else 0))
    return int(L.ubs_rasterize_workspace_bytes(C, N, 1 if backward else 0))


def prepacked_blend_workspace(C: int, N: int, channels: int, tile_size: int, tile_width: int, tile_height: int,
                              device):
    """A blend workspace for the fused projection to write the blend's
    records into (beta_project(records=...)), then handed to
    rasterize_to_pixels(prepacked_workspace=...); None where the blend packs
    them itself (not 3 channels, 32-pixel tiles, transmittance checkpoints,
    packed records off)."""
    if channels != 3 or tile_size not in (8, 16) or _t_checkpoint:
        return None
    return _blend_workspace(C, N, False, device, tile_width, tile_height)[0]


def _blend_workspace(C: int, N: int, backward: bool, device, tile_width: int, tile_height: int):
    if not _packed_records or C * N == 0:
        return None, 0
    nbytes = blend_workspace_bytes(C, N, tile_width, tile_height, backward)
    return torch.empty((nbytes,), dtype=torch.uint8, device=device), nbytes


def raster_fwd_raw(means2d, conics, colors, opacities, betas, backgrounds, masks, width, height, tile_size,
                   isect_offsets, flatten_ids):
    """Blend forward: (render_colors, render_alphas, last_ids, T_final)."""
    return _raster_fwd(means2d, conics, colors, opacities, betas, backgrounds, masks, width, height, tile_size,
                       isect_offsets, flatten_ids, False)[:4]


def raster_fwd_ckpt_raw(means2d, conics, colors, opacities, betas, backgrounds, masks, width, height, tile_size,
                        isect_offsets, flatten_ids):
    """Blend forward that also stores the transmittance checkpoints:
    (render_colors, render_alphas, last_ids, T_final, T_ckpt)."""
    return _raster_fwd(means2d, conics, colors, opacities, betas, backgrounds, masks, width, height, tile_size,
                       isect_offsets, flatten_ids, True)[:5]


def _raster_fwd(means2d, conics, colors, opacities, betas, backgrounds, masks, width, height, tile_size,
                isect_offsets, flatten_ids, checkpoint: bool, prepacked=None):
    """(render_colors, render_alphas, last_ids, T_final, T_ckpt or None, the
    workspace: None, or the packed records and, with the tile order on, the
    order, which the backward of the same inputs can reuse)."""
    require_cuda(means2d, conics, colors, opacities, betas, backgrounds, isect_offsets, flatten_ids)
    require_dtype(_F32, means2d, conics, colors, opacities, betas, backgrounds)
    require_dtype(_I32, isect_offsets, flatten_ids)
    C, N = opacities.shape
    channels = colors.shape[-1]
    th, tw = isect_offsets.shape[1:3]
    dev = means2d.device
    render_colors = torch.empty((C, height, width, channels), dtype=torch.float32, device=dev)
    render_alphas = torch.empty((C, height, width, 1), dtype=torch.float32, device=dev)
    last_ids = torch.empty((C, height, width), dtype=torch.int32, device=dev)
    T_final = torch.empty((C, height, width), dtype=torch.float32, device=dev)
    mp = _mask_ptr(masks)
    if prepacked is not None and not checkpoint:  # records written by the fused projection
        ws, ws_bytes = prepacked, prepacked.numel()
    else:
        prepacked = None
        ws, ws_bytes = _blend_workspace(C, N, False, dev, tw, th)
    if not checkpoint:
        entry = "ubs_rasterize_to_pixels_fwd_prepacked" if prepacked is not None else "ubs_rasterize_to_pixels_fwd"
        _call(entry, C, N, flatten_ids.numel(), channels, ptr(means2d), ptr(conics),
              ptr(colors), ptr(opacities), ptr(betas), ptr(backgrounds), mp[0] if mp else None, width, height,
              tile_size, tw, th, ptr(isect_offsets), ptr(flatten_ids), ptr(render_colors), ptr(render_alphas),
              ptr(last_ids), ptr(T_final), ptr(ws), ws_bytes, stream_of(means2d))
        return render_colors, render_alphas, last_ids, T_final, None, ws
    n_ck = int(lib().ubs_transmittance_checkpoint_floats(C, flatten_ids.numel(), tw, th, tile_size))
    T_ckpt = torch.empty((n_ck,), dtype=torch.float32, device=dev)
    _call("ubs_rasterize_to_pixels_fwd_ckpt", C, N, flatten_ids.numel(), channels, ptr(means2d), ptr(conics),
          ptr(colors), ptr(opacities), ptr(betas), ptr(backgrounds), mp[0] if mp else None, width, height,
          tile_size, tw, th, ptr(isect_offsets), ptr(flatten_ids), ptr(render_colors), ptr(render_alphas),
          ptr(last_ids), ptr(T_final), ptr(T_ckpt), n_ck, ptr(ws), ws_bytes, stream_of(means2d))
    return render_colors, render_alphas, last_ids, T_final, T_ckpt, ws


def raster_bwd_raw(means2d, conics, colors, opacities, betas, backgrounds, masks, width, height, tile_size,
                   isect_offsets, flatten_ids, render_alphas, last_ids, T_final, v_render_colors,
                   v_render_alphas, T_ckpt=None, fwd_workspace=None):
    """Blend backward.  fwd_workspace: the workspace the forward of the same
    inputs returned (records and tile order, sized with the order): its
    records and order are reused (ubs_rasterize_to_pixels_bwd_reuse) instead
    of packed and computed again."""
    require_cuda(means2d, conics, colors, opacities, betas, backgrounds, isect_offsets, flatten_ids,
                 render_alphas, last_ids, T_final, v_render_colors, v_render_alphas)
    require_dtype(_F32, means2d, conics, colors, opacities, betas, backgrounds, render_alphas, T_final,
                  v_render_colors, v_render_alphas)
    require_dtype(_I32, isect_offsets, flatten_ids, last_ids)
    C, N = opacities.shape  # (v_render_alphas None: a zero gradient)
    channels = colors.shape[-1]
    th, tw = isect_offsets.shape[1:3]
    v_means2d = torch.empty_like(means2d)
    v_conics = torch.empty_like(conics)
    v_colors = torch.empty_like(colors)
    v_opacities = torch.empty_like(opacities)
    v_betas = torch.empty_like(betas)
    mp = _mask_ptr(masks)
    if fwd_workspace is not None and _packed_records and C * N > 0:
        if T_ckpt is not None:
            require_cuda(T_ckpt)
            require_dtype(_F32, T_ckpt)
        gb = int(lib().ubs_rasterize_workspace_bytes(C, N, 0))
        gws = torch.empty((gb,), dtype=torch.uint8, device=means2d.device)
        _call("ubs_rasterize_to_pixels_bwd_reuse", C, N, flatten_ids.numel(), channels, ptr(means2d), ptr(conics),
              ptr(colors), ptr(opacities), ptr(betas), ptr(backgrounds), mp[0] if mp else None, width, height,
              tile_size, tw, th, ptr(isect_offsets), ptr(flatten_ids), ptr(render_alphas), ptr(last_ids),
              ptr(T_final), ptr(T_ckpt), 0 if T_ckpt is None else T_ckpt.numel(), ptr(v_render_colors),
              ptr(v_render_alphas), ptr(v_means2d), ptr(v_conics), ptr(v_colors), ptr(v_opacities), ptr(v_betas),
              ptr(fwd_workspace), fwd_workspace.numel(), ptr(gws), gb, stream_of(means2d))
        return v_means2d, v_conics, v_colors, v_opacities, v_betas
    ws, ws_bytes = _blend_workspace(C, N, True, means2d.device, tw, th)
    if T_ckpt is None:
        _call("ubs_rasterize_to_pixels_bwd", C, N, flatten_ids.numel(), channels, ptr(means2d), ptr(conics),
              ptr(colors), ptr(opacities), ptr(betas), ptr(backgrounds), mp[0] if mp else None, width, height,
              tile_size, tw, th, ptr(isect_offsets), ptr(flatten_ids), ptr(render_alphas), ptr(last_ids),
              ptr(T_final), ptr(v_render_colors), ptr(v_render_alphas), ptr(v_means2d), ptr(v_conics),
              ptr(v_colors), ptr(v_opacities), ptr(v_betas), ptr(ws), ws_bytes, stream_of(means2d))
    else:
        require_cuda(T_ckpt)
        require_dtype(_F32, T_ckpt)
        _call("ubs_rasterize_to_pixels_bwd_ckpt", C, N, flatten_ids.numel(), channels, ptr(means2d), ptr(conics),
              ptr(colors), ptr(opacities), ptr(betas), ptr(backgrounds), mp[0] if mp else None, width, height,
              tile_size, tw, th, ptr(isect_offsets), ptr(flatten_ids), ptr(render_alphas), ptr(last_ids),
              ptr(T_final), ptr(T_ckpt), T_ckpt.numel(), ptr(v_render_colors), ptr(v_render_alphas),
              ptr(v_means2d), ptr(v_conics), ptr(v_colors), ptr(v_opacities), ptr(v_betas), ptr(ws), ws_bytes,
              stream_of(means2d))
    return v_means2d, v_conics, v_colors, v_opacities, v_betas


def activate_fwd_raw(xyz, mean_rest, scale_raw, opacity_raw, beta_raw, quats, cam_pos, timestamp):
    """ubs_beta_activate_fwd: (mean, scale, opacity, beta, query or None, rot or None)."""
    require_cuda(xyz, mean_rest, scale_raw, opacity_raw, beta_raw, quats, cam_pos)
    require_dtype(_F32, xyz, mean_rest, scale_raw, opacity_raw, beta_raw, quats, cam_pos)
    N, D = scale_raw.shape
    assert xyz.shape == (N, 3) and mean_rest.shape == (N, D - 3), (xyz.shape, mean_rest.shape)
    assert opacity_raw.shape == (N, 1) and beta_raw.shape == (N, D - 2), (opacity_raw.shape, beta_raw.shape)
    assert quats is None or quats.shape == (N, 4), quats.shape
    assert cam_pos is None or cam_pos.numel() == 3
    dev, dt = xyz.device, xyz.dtype
    mean = torch.empty((N, D), dtype=dt, device=dev)
    scale = torch.empty((N, D), dtype=dt, device=dev)
    opacity = torch.empty((N, 1), dtype=dt, device=dev)
    beta = torch.empty((N, D - 2), dtype=dt, device=dev)
    query = torch.empty((N, D - 3), dtype=dt, device=dev) if cam_pos is not None else None
    rot = torch.empty((N, 3, 3), dtype=dt, device=dev) if quats is not None else None
    _call("ubs_beta_activate_fwd", N, D, ptr(xyz), ptr(mean_rest), ptr(scale_raw), ptr(opacity_raw),
          ptr(beta_raw), ptr(quats), ptr(cam_pos), float(timestamp), ptr(mean), ptr(scale), ptr(opacity),
          ptr(beta), ptr(query), ptr(rot), stream_of(xyz))
    return mean, scale, opacity, beta, query, rot


def activate_bwd_raw(scale_raw, opacity_raw, beta_raw, quats, v_mean, v_scale, v_opacity, v_beta, v_rot):
    """ubs_beta_activate_bwd: (v_xyz, v_mean_rest, v_scale_raw, v_opacity_raw, v_beta_raw, v_quats or None)."""
    grads = [None if g is None else g.contiguous() for g in (v_mean, v_scale, v_opacity, v_beta, v_rot)]
    require_cuda(scale_raw, opacity_raw, beta_raw, quats, *grads)
    require_dtype(_F32, scale_raw, opacity_raw, beta_raw, quats, *grads)
    N, D = scale_raw.shape
    dev, dt = scale_raw.device, scale_raw.dtype
    v_xyz = torch.empty((N, 3), dtype=dt, device=dev)
    v_mean_rest = torch.empty((N, D - 3), dtype=dt, device=dev)
    v_scale_raw = torch.empty_like(scale_raw)
    v_opacity_raw = torch.empty_like(opacity_raw)
    v_beta_raw = torch.empty_like(beta_raw)
    v_quats = torch.empty_like(quats) if quats is not None else None
    _call("ubs_beta_activate_bwd", N, D, ptr(scale_raw), ptr(opacity_raw), ptr(beta_raw), ptr(quats),
          *[ptr(g) for g in grads], ptr(v_xyz), ptr(v_mean_rest), ptr(v_scale_raw), ptr(v_opacity_raw),
          ptr(v_beta_raw), ptr(v_quats), stream_of(scale_raw))
    return v_xyz, v_mean_rest, v_scale_raw, v_opacity_raw, v_beta_raw, v_quats


def cond_fwd_raw(means, covars, opacities, betas, query):
    """cond_mean_convariance_opacity_fwd: (means [N,3], covars [N,3,3], opacities [N,1])."""
    require_cuda(means, covars, opacities, betas, query)
    require_dtype(_F32, means, covars, opacities, betas, query)
    assert means.dim() == 2, "means must be [N,D]"
    assert covars.dim() == 3, "covars must be [N,D,D]"
    assert opacities.dim() == 2 and opacities.size(1) == 1, "opacities must be [N,1]"
    assert betas.dim() == 2 and query.dim() == 2, "betas/query must be [N,C]"
    N, D = means.shape
    assert covars.shape == (N, D, D), "covars must match [N,D,D]"
    assert query.size(0) == N and betas.size(0) == N, "batch sizes must match"
    assert D >= 4, "D must be >= 4"
    assert D - 3 == betas.size(1) == query.size(1), "C must be D-3"
    om = torch.empty((N, 3), dtype=means.dtype, device=means.device)
    oV = torch.empty((N, 3, 3), dtype=means.dtype, device=means.device)
    oo = torch.empty((N, 1), dtype=means.dtype, device=means.device)
    _call("ubs_cond_mean_convariance_opacity_fwd", N, D, ptr(means), ptr(covars), ptr(opacities),
          ptr(betas), ptr(query), ptr(om), ptr(oV), ptr(oo), stream_of(means))
    return om, oV, oo


def cond_bwd_raw(means, covars, opacities, betas, query, v_m, v_V, v_o):
    """cond_mean_convariance_opacity_bwd: (v_means [N,D], v_covars [N,D,D], v_opacities [N,1], v_betas [N,C])."""
    require_cuda(means, covars, opacities, betas, query, v_m, v_V, v_o)
    require_dtype(_F32, means, covars, opacities, betas, query, v_m, v_V, v_o)
    N, D = means.shape
    v_m, v_V, v_o = v_m.contiguous(), v_V.contiguous(), v_o.contiguous()
    gm = torch.empty_like(means)
    gV = torch.empty_like(covars)
    go = torch.empty_like(opacities)
    gb = torch.empty_like(betas)
    _call("ubs_cond_mean_convariance_opacity_bwd", N, D, ptr(means), ptr(covars), ptr(opacities),
          ptr(betas), ptr(query), ptr(v_m), ptr(v_V), ptr(v_o), ptr(gm), ptr(gV), ptr(go), ptr(gb),
          stream_of(means))
    return gm, gV, go, gb


def rotmat_fwd_raw(l_triangle):
    """l_triangle_to_rotmat_fwd: [N,3] -> [N,3,3]."""
    require_cuda(l_triangle)
    require_dtype(_F32, l_triangle)
    N = l_triangle.size(0)
    R = torch.empty((N, 3, 3), dtype=l_triangle.dtype, device=l_triangle.device)
    _call("ubs_l_triangle_to_rotmat_fwd", N, ptr(l_triangle), ptr(R), stream_of(l_triangle))
    return R


def rotmat_bwd_raw(l_triangle, v_R):
    """l_triangle_to_rotmat_bwd: v_R [N,3,3] -> v_l_triangle [N,3]."""
    require_cuda(l_triangle, v_R)
    require_dtype(_F32, l_triangle, v_R)
    v_R = v_R.contiguous()
    out = torch.empty_like(l_triangle)
    _call("ubs_l_triangle_to_rotmat_bwd", l_triangle.size(0), ptr(v_R), ptr(out), stream_of(v_R))
    return out


def covar_fwd_raw(rot, scale, l_triangle, rest_i, rest_j, spatial_block):
    """rot_scale_l_triangle_to_covar_fwd: [N,D,D] (or [N,3,3] for the spatial block)."""
    require_cuda(rot, scale, l_triangle, rest_i, rest_j)
    require_dtype(_F32, rot, scale, l_triangle)
    N, D = scale.shape
    rest_i = rest_i.to(torch.int32).contiguous()
    rest_j = rest_j.to(torch.int32).contiguous()
    dim = 3 if (spatial_block or D == 3) else D
    covar = torch.empty((N, dim, dim), dtype=rot.dtype, device=rot.device)
    _call("ubs_rot_scale_l_triangle_to_covar_fwd", N, D, ptr(rot), ptr(scale), ptr(l_triangle),
          ptr(rest_i), ptr(rest_j), rest_i.numel(), int(spatial_block), ptr(covar), stream_of(rot))
    return covar


def covar_bwd_raw(rot, scale, l_triangle, rest_i, rest_j, spatial_block, v_covar):
    """rot_scale_l_triangle_to_covar_bwd: (v_rot [N,3,3], v_scale [N,D], v_l_triangle [N,D(D-1)/2])."""
    require_cuda(rot, scale, l_triangle, rest_i, rest_j, v_covar)
    require_dtype(_F32, rot, scale, l_triangle, v_covar)
    N, D = scale.shape
    rest_i = rest_i.to(torch.int32).contiguous()
    rest_j = rest_j.to(torch.int32).contiguous()
    v_covar = v_covar.contiguous()
    v_rot = torch.empty_like(rot)
    v_scale = torch.empty_like(scale)
    v_l = torch.empty_like(l_triangle)
    _call("ubs_rot_scale_l_triangle_to_covar_bwd", N, D, ptr(rot), ptr(scale), ptr(l_triangle),
          ptr(rest_i), ptr(rest_j), rest_i.numel(), int(spatial_block), ptr(v_covar), ptr(v_rot),
          ptr(v_scale), ptr(v_l), stream_of(rot))
    return v_rot, v_scale, v_l


def projection_fwd_raw(means, covars, quats, scales, viewmats, Ks, width, height, eps2d, near_plane, far_plane,
                       radius_clip, calc_compensations, ortho, covar_layout=0):
    """fully_fused_projection_fwd: (radii, means2d, depths, conics, compensations or None)."""
    require_cuda(means, covars, quats, scales, viewmats, Ks)
    require_dtype(_F32, means, covars, quats, scales, viewmats, Ks)
    C, N = viewmats.size(0), means.size(0)
    dev = means.device
    radii = torch.empty((C, N), dtype=torch.int32, device=dev)
    means2d = torch.empty((C, N, 2), dtype=means.dtype, device=dev)
    depths = torch.empty((C, N), dtype=means.dtype, device=dev)
    conics = torch.empty((C, N, 3), dtype=means.dtype, device=dev)
    comp = torch.zeros((C, N), dtype=means.dtype, device=dev) if calc_compensations else None
    _call("ubs_fully_fused_projection_fwd", C, N, ptr(means), ptr(covars), covar_layout, ptr(quats),
          ptr(scales), ptr(viewmats), ptr(Ks), width, height, eps2d, near_plane, far_plane, radius_clip,
          int(ortho), ptr(radii), ptr(means2d), ptr(depths), ptr(conics), ptr(comp), stream_of(means))
    return radii, means2d, depths, conics, comp


def projection_bwd_raw(means, covars, quats, scales, viewmats, Ks, width, height, eps2d, ortho, radii, conics,
                       compensations, v_means2d, v_depths, v_conics, v_compensations, viewmats_requires_grad,
                       covar_layout=0):
    """fully_fused_projection_bwd: (v_means, v_covars, v_quats, v_scales, v_viewmats); the
    entries of the absent covariance parametrisation and v_viewmats when not required are None."""
    require_cuda(means, covars, quats, scales, viewmats, Ks, radii, conics, v_means2d, v_depths, v_conics)
    require_dtype(_F32, means, covars, quats, scales, viewmats, Ks, conics, compensations, v_means2d, v_depths,
                  v_conics, v_compensations)
    require_dtype(_I32, radii)
    C, N = viewmats.size(0), means.size(0)
    v_means = torch.empty_like(means)
    v_covars = torch.empty_like(covars) if covars is not None else None
    v_quats = torch.empty_like(quats) if covars is None else None
    v_scales = torch.empty_like(scales) if covars is None else None
    v_viewmats = torch.empty_like(viewmats) if viewmats_requires_grad else None
    v_means2d = v_means2d.contiguous()
    v_depths = v_depths.contiguous()
    v_conics = v_conics.contiguous()
    if v_compensations is not None:
        v_compensations = v_compensations.contiguous()
    _call("ubs_fully_fused_projection_bwd", C, N, ptr(means), ptr(covars), covar_layout, ptr(quats),
          ptr(scales), ptr(viewmats), ptr(Ks), width, height, eps2d, int(ortho), ptr(radii),
          ptr(conics), ptr(compensations), ptr(v_means2d), ptr(v_depths), ptr(v_conics),
          ptr(v_compensations) if compensations is not None else None, ptr(v_means), ptr(v_covars),
          ptr(v_quats), ptr(v_scales), ptr(v_viewmats), stream_of(means))
    return v_means, v_covars, v_quats, v_scales, v_viewmats


def beta_project_fwd_raw(xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats, cam_pos, timestamps,
                         viewmats, Ks, width, height, image_wh, eps2d, near_plane, far_plane, radius_clip, ortho,
                         rgb=None, records=None):
    """ubs_beta_project_fwd: (radii, means2d, depths, conics, opacities, betas), each [C,N,...]."""
    require_cuda(xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats, cam_pos, timestamps,
                 viewmats, Ks, image_wh)
    require_dtype(_F32, xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats, cam_pos, timestamps,
                  viewmats, Ks)
    require_dtype(_I32, image_wh)
    N, D = scale_raw.shape
    C = viewmats.size(0)
    assert xyz.shape == (N, 3) and mean_rest.shape == (N, D - 3), (xyz.shape, mean_rest.shape)
    assert l_triangle.shape == (N, D * (D - 1) // 2), l_triangle.shape
    assert opacity_raw.shape == (N, 1) and beta_raw.shape == (N, D - 2), (opacity_raw.shape, beta_raw.shape)
    assert quats is None or quats.shape == (N, 4), quats.shape
    assert viewmats.shape == (C, 4, 4) and Ks.shape == (C, 3, 3), (viewmats.shape, Ks.shape)
    assert cam_pos is None or cam_pos.shape == (C, 3), cam_pos.shape
    assert timestamps is None or timestamps.shape == (C,), timestamps.shape
    assert image_wh is None or image_wh.shape == (C, 2), image_wh.shape
    dev, dt = xyz.device, xyz.dtype
    radii = torch.empty((C, N), dtype=torch.int32, device=dev)
    means2d = torch.empty((C, N, 2), dtype=dt, device=dev)
    depths = torch.empty((C, N), dtype=dt, device=dev)
    conics = torch.empty((C, N, 3), dtype=dt, device=dev)
    opacities = torch.empty((C, N), dtype=dt, device=dev)
    betas = torch.empty((C, N), dtype=dt, device=dev)
    args = (N, D, C, ptr(xyz), ptr(mean_rest), ptr(scale_raw), ptr(l_triangle), ptr(opacity_raw), ptr(beta_raw),
            ptr(quats), ptr(cam_pos), ptr(timestamps), ptr(viewmats), ptr(Ks), width, height, ptr(image_wh), eps2d,
            near_plane, far_plane, radius_clip, int(ortho), ptr(radii), ptr(means2d), ptr(depths), ptr(conics),
            ptr(opacities), ptr(betas))
    if records is not None:  # the blend's records too (rgb [N,3])
        require_cuda(rgb, records)
        require_dtype(_F32, rgb)
        assert rgb.shape == (N, 3), rgb.shape
        _call("ubs_beta_project_fwd_rec", *args, ptr(rgb), ptr(records), stream_of(xyz))
    else:
        _call("ubs_beta_project_fwd", *args, stream_of(xyz))
    return radii, means2d, depths, conics, opacities, betas


def beta_project_bwd_raw(xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats, cam_pos, timestamps,
                         viewmats, Ks, width, height, image_wh, eps2d, ortho, radii, conics, v_means2d, v_depths,
                         v_conics, v_opacities, v_betas):
    """ubs_beta_project_bwd: (v_xyz, v_mean_rest, v_scale_raw, v_l_triangle, v_opacity_raw, v_beta_raw,
    v_quats or None)."""
    grads = [None if g is None else g.contiguous() for g in (v_means2d, v_depths, v_conics, v_opacities, v_betas)]
    require_cuda(xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats, cam_pos, timestamps,
                 viewmats, Ks, image_wh, radii, conics, *grads)
    require_dtype(_F32, xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats, cam_pos, timestamps,
                  viewmats, Ks, conics, *grads)
    require_dtype(_I32, image_wh, radii)
    N, D = scale_raw.shape
    C = viewmats.size(0)
    v_xyz = torch.empty_like(xyz)
    v_mean_rest = torch.empty_like(mean_rest)
    v_scale_raw = torch.empty_like(scale_raw)
    v_l_triangle = torch.empty_like(l_triangle)
    v_opacity_raw = torch.empty_like(opacity_raw)
    v_beta_raw = torch.empty_like(beta_raw)
    v_quats = torch.empty_like(quats) if quats is not None else None
    _call("ubs_beta_project_bwd", N, D, C, ptr(xyz), ptr(mean_rest), ptr(scale_raw), ptr(l_triangle),
          ptr(opacity_raw), ptr(beta_raw), ptr(quats), ptr(cam_pos), ptr(timestamps), ptr(viewmats), ptr(Ks),
          width, height, ptr(image_wh), eps2d, int(ortho), ptr(radii), ptr(conics), *[ptr(g) for g in grads],
          ptr(v_xyz), ptr(v_mean_rest), ptr(v_scale_raw), ptr(v_l_triangle), ptr(v_opacity_raw), ptr(v_beta_raw),
          ptr(v_quats), stream_of(xyz))
    return v_xyz, v_mean_rest, v_scale_raw, v_l_triangle, v_opacity_raw, v_beta_raw, v_quats


# ----------------------------------------------------------------------------
# autograd Functions (reference: _wrapper.py:573-684, 807-1050)
# ----------------------------------------------------------------------------
class _BetaProject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats, viewmats, Ks, cam_pos,
                timestamps, image_wh, width, height, eps2d, near_plane, far_plane, radius_clip, ortho, rgb=None,
                records=None):
        radii, means2d, depths, conics, opacities, betas = beta_project_fwd_raw(
            xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats, cam_pos, timestamps, viewmats, Ks,
            width, height, image_wh, eps2d, near_plane, far_plane, radius_clip, ortho, rgb, records)
        ctx.save_for_backward(xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats, viewmats, Ks,
                              cam_pos, timestamps, image_wh, radii, conics)
        ctx.width, ctx.height, ctx.eps2d, ctx.ortho = width, height, eps2d, ortho
        ctx.mark_non_differentiable(radii)
        ctx.set_materialize_grads(False)  # the radii and depths gradients come back as None
        return radii, means2d, depths, conics, opacities, betas

    @staticmethod
    def backward(ctx, v_radii, v_means2d, v_depths, v_conics, v_opacities, v_betas):
        (xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats, viewmats, Ks, cam_pos, timestamps,
         image_wh, radii, conics) = ctx.saved_tensors
        C, N = radii.shape

        def g(t, shape):
            return torch.zeros(shape, dtype=xyz.dtype, device=xyz.device) if t is None else t

        grads = beta_project_bwd_raw(
            xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats, cam_pos, timestamps, viewmats, Ks,
            ctx.width, ctx.height, image_wh, ctx.eps2d, ctx.ortho, radii, conics, g(v_means2d, (C, N, 2)),
            v_depths, g(v_conics, (C, N, 3)), g(v_opacities, (C, N)), g(v_betas, (C, N)))  # (None: zeros)
        out = [gr if need else None for gr, need in zip(grads, ctx.needs_input_grad[:7])]
        return (*out, None, None, None, None, None, None, None, None, None, None, None, None, None, None)


class _BetaActivations(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, mean_rest, scale_raw, opacity_raw, beta_raw, quats, cam_pos, timestamp):
        mean, scale, opacity, beta, query, rot = activate_fwd_raw(
            xyz, mean_rest, scale_raw, opacity_raw, beta_raw, quats, cam_pos, timestamp)
        ctx.save_for_backward(scale_raw, opacity_raw, beta_raw, quats)
        if query is not None:
            ctx.mark_non_differentiable(query)
        return mean, scale, opacity, beta, query, rot

    @staticmethod
    def backward(ctx, v_mean, v_scale, v_opacity, v_beta, v_query, v_rot):
        scale_raw, opacity_raw, beta_raw, quats = ctx.saved_tensors
        if quats is not None and not ctx.needs_input_grad[5]:
            quats = None
        gx, gm, gs, go, gb, gq = activate_bwd_raw(scale_raw, opacity_raw, beta_raw, quats, v_mean, v_scale,
                                                  v_opacity, v_beta, v_rot)
        return gx, gm, gs, go, gb, gq, None, None


class _CondMeanConvarianceOpacity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means, covars, opacities, betas, query):
        om, oV, oo = cond_fwd_raw(means, covars, opacities, betas, query)
        ctx.save_for_backward(means, covars, opacities, betas, query)
        return om, oV, oo

    @staticmethod
    def backward(ctx, v_m, v_V, v_o):
        means, covars, opacities, betas, query = ctx.saved_tensors
        gm, gV, go, gb = cond_bwd_raw(means, covars, opacities, betas, query, v_m, v_V, v_o)
        return gm, gV, go, gb, None


class _LTriangleToRotmat(torch.autograd.Function):
    @staticmethod
    def forward(ctx, l_triangle):
        R = rotmat_fwd_raw(l_triangle)
        ctx.save_for_backward(l_triangle)
        return R

    @staticmethod
    def backward(ctx, v_R):
        (l_triangle,) = ctx.saved_tensors
        return rotmat_bwd_raw(l_triangle, v_R)


class _RotScaleLTriangleToCovar(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rot, scale, l_triangle, rest_i, rest_j, spatial_block):
        covar = covar_fwd_raw(rot, scale, l_triangle, rest_i, rest_j, spatial_block)
        ctx.save_for_backward(rot, scale, l_triangle, rest_i, rest_j)
        ctx.spatial_block = spatial_block
        return covar

    @staticmethod
    def backward(ctx, v_covar):
        rot, scale, l_triangle, rest_i, rest_j = ctx.saved_tensors
        if not any(ctx.needs_input_grad[:3]):
            return (None, None, None, None, None, None)
        v_rot, v_scale, v_l = covar_bwd_raw(rot, scale, l_triangle, rest_i, rest_j, ctx.spatial_block, v_covar)
        return v_rot, v_scale, v_l, None, None, None


class _FullyFusedProjection(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means, covars, quats, scales, viewmats, Ks, width, height, eps2d, near_plane,
                far_plane, radius_clip, calc_compensations, ortho, covar_layout):
        radii, means2d, depths, conics, comp = projection_fwd_raw(
            means, covars, quats, scales, viewmats, Ks, width, height, eps2d, near_plane, far_plane, radius_clip,
            calc_compensations, ortho, covar_layout)
        ctx.save_for_backward(means, covars, quats, scales, viewmats, Ks, radii, conics, comp)
        ctx.width, ctx.height, ctx.eps2d, ctx.ortho = width, height, eps2d, ortho
        ctx.covar_layout = covar_layout
        return radii, means2d, depths, conics, comp

    @staticmethod
    def backward(ctx, v_radii, v_means2d, v_depths, v_conics, v_compensations):
        means, covars, quats, scales, viewmats, Ks, radii, conics, comp = ctx.saved_tensors
        v_means, v_covars, v_quats, v_scales, v_viewmats = projection_bwd_raw(
            means, covars, quats, scales, viewmats, Ks, ctx.width, ctx.height, ctx.eps2d, ctx.ortho, radii, conics,
            comp, v_means2d, v_depths, v_conics, v_compensations, ctx.needs_input_grad[4], ctx.covar_layout)
        if not ctx.needs_input_grad[0]:
            v_means = None
        if not ctx.needs_input_grad[1]:
            v_covars = None
        if not ctx.needs_input_grad[2]:
            v_quats = None
        if not ctx.needs_input_grad[3]:
            v_scales = None
        return (v_means, v_covars, v_quats, v_scales, v_viewmats,
                None, None, None, None, None, None, None, None, None, None)


class _RasterizeToPixels(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means2d, conics, colors, opacities, betas, backgrounds, masks, width, height, tile_size,
                isect_offsets, flatten_ids, prepacked=None):
        ckpt = _t_checkpoint and any(ctx.needs_input_grad[:5])
        # an output without a gradient comes back as None, not a filled zero tensor
        ctx.set_materialize_grads(False)
        render_colors, render_alphas, last_ids, T_final, T_ckpt, ws = _raster_fwd(
            means2d, conics, colors, opacities, betas, backgrounds, masks, width, height, tile_size,
            isect_offsets, flatten_ids, ckpt, prepacked)
        ctx.save_for_backward(means2d, conics, colors, opacities, betas, backgrounds, masks, isect_offsets,
                              flatten_ids, render_alphas, last_ids, T_final, T_ckpt)
        ctx.width, ctx.height, ctx.tile_size = width, height, tile_size
        # the forward's packed records and tile order, for the backward to reuse
        ctx.fwd_ws = ws if (ws is not None and _tile_order and any(ctx.needs_input_grad[:5])) else None
        return render_colors, render_alphas

    @staticmethod
    def backward(ctx, v_render_colors, v_render_alphas):
        (means2d, conics, colors, opacities, betas, backgrounds, masks, isect_offsets, flatten_ids,
         render_alphas, last_ids, T_final, T_ckpt) = ctx.saved_tensors
        if v_render_colors is None:
            v_render_colors = torch.zeros((*render_alphas.shape[:3], colors.shape[-1]), dtype=torch.float32,
                                          device=render_alphas.device)
        v_means2d, v_conics, v_colors, v_opacities, v_betas = raster_bwd_raw(
            means2d, conics, colors, opacities, betas, backgrounds, masks, ctx.width, ctx.height,
            ctx.tile_size, isect_offsets, flatten_ids, render_alphas, last_ids, T_final,
            v_render_colors.contiguous(), None if v_render_alphas is None else v_render_alphas.contiguous(),
            T_ckpt=T_ckpt, fwd_workspace=ctx.fwd_ws)
        ctx.fwd_ws = None
        if ctx.needs_input_grad[5]:
            v_backgrounds = (v_render_colors * (1.0 - render_alphas).float()).sum(dim=(1, 2))
        else:
            v_backgrounds = None
        return (v_means2d, v_conics, v_colors, v_opacities, v_betas, v_backgrounds,
                None, None, None, None, None, None, None)
