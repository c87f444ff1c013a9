"""Fixed-shape forward of a batch of views of resident scenes, with no host
synchronisation, captured once as a hipGraph per (image size, slots) and
replayed for every batch of that shape.

The reference renders one view per request (BetaModel.view,
scene/beta_model.py:725-831) and reads the intersection count on the host
before it can size the lists (isect_tiles.cu:180-181).  ``ViewBatchGraph``
instead sizes the lists for a capacity on the host ahead of time and lets
the count stay on the device:

  1. ubs_beta_project_fwd_table: every slot's view of its own scene (a scene
     table in device memory), rows of ``row_stride`` entries, colours too;
  2. ubs_isect_tiles_count_range + depth keys + depth sort + ordered scan:
     tile counts and the depth order, the count in ``info`` on the device;
  3. ubs_isect_tiles_emit_ordered_cap: the entries up to the capacity;
  4. ubs_radix_sort_pairs_dev on the camera/tile bits, then
     ubs_isect_offset_encode_dev and ubs_rasterize_to_pixels_fwd_dev.

Every launch has a fixed shape, so the sequence is captured in one graph;
a replay renders whatever the static inputs (scene table, cameras, times)
hold.  The count is read after the work finished: above the capacity the
batch overflowed (its images are invalid) and the caller renders it again
through the host-sized path and grows the capacity.  Empty slots (N = 0)
render the background.  Per view the images equal render_views' bit for
bit: the same device functions, lists and blend.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Sequence

import torch
from torch import Tensor

from .beta import BetaParams
from .cuda._backend import lib
from .cuda._wrapper import _call, blend_workspace_bytes, camera_centers, ptr, stream_of


class _ViewScene(ctypes.Structure):
    """include/ubs_capi.h: ubs_view_scene."""
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("xyz", "mean_rest", "scale_raw", "l_triangle", "opacity_raw", "beta_raw", "quats", "rgb")] + \
               [("N", ctypes.c_int64), ("D", ctypes.c_int64)]


def scene_entry(p: Optional[BetaParams]) -> _ViewScene:
    """The table entry of a resident scene (None: an empty slot)."""
    e = _ViewScene()
    if p is None:
        e.N, e.D = 0, 6
        return e
    for name in ("xyz", "mean_rest", "scale_raw", "l_triangle", "opacity_raw", "beta_raw", "rgb"):
        t = getattr(p, name)
        assert t.is_contiguous() and t.dtype == torch.float32, name
        setattr(e, name, t.data_ptr())
    e.quats = p.quats.data_ptr() if p.quats is not None else None
    e.N, e.D = p.xyz.shape[0], p.D
    return e


class ViewBatchGraph:
    """S slots of one image size (W, H) over scenes with at most
    ``row_stride`` primitives, intersection lists for ``cap`` entries.
    ``set_views`` fills the static inputs, ``run`` enqueues the launches
    (eagerly, or inside a capture), ``capture`` records them as a graph and
    ``replay`` re-runs it; ``images`` / ``alphas`` hold the last batch."""

    def __init__(self, device, slots: int, width: int, height: int, row_stride: int, cap: int,
                 tile_size: int = 16, eps2d: float = 0.3, near_plane: float = 0.01, far_plane: float = 1e10,
                 radius_clip: float = 0.0):
        self.device = torch.device(device)
        self.S, self.W, self.H, self.R, self.cap = int(slots), int(width), int(height), int(row_stride), int(cap)
        assert self.S >= 1 and self.R >= 1 and self.cap >= 1
        assert self.S * self.R < (1 << 31) and self.cap < (1 << 31)
        self.ts = tile_size
        self.tw, self.th = math.ceil(self.W / float(tile_size)), math.ceil(self.H / float(tile_size))
        self.eps2d, self.near, self.far, self.rclip = eps2d, near_plane, far_plane, radius_clip
        dev, S = self.device, self.S
        f32 = dict(dtype=torch.float32, device=dev)
        # static inputs
        self.table = torch.zeros((S * ctypes.sizeof(_ViewScene),), dtype=torch.uint8, device=dev)
        self.viewmats = torch.zeros((S, 4, 4), **f32)
        self.Ks = torch.zeros((S, 3, 3), **f32)
        self.cam_pos = torch.zeros((S, 3), **f32)
        self.times = torch.zeros((S,), **f32)
        self.wh = torch.tensor([[self.W, self.H]] * S, dtype=torch.int32, device=dev)
        # staging for the inputs (pinned host memory, copied on the batch's stream)
        self._h_table = torch.zeros((S * ctypes.sizeof(_ViewScene),), dtype=torch.uint8, pin_memory=True)
        self._h_cams = torch.zeros((S, 16 + 9 + 3 + 1), dtype=torch.float32, pin_memory=True)
        # outputs of the last batch
        self.images = torch.zeros((S, self.H, self.W, 3), **f32)
        self.alphas = torch.zeros((S, self.H, self.W, 1), **f32)
        self.info = torch.zeros((3,), dtype=torch.int64, device=dev)
        self.graph: Optional[torch.cuda.CUDAGraph] = None
        L = lib()
        self.end_bit = 32 + int(L.ubs_isect_prefix_bits(S, self.tw, self.th))
        self.views: List = []

    # -- inputs ----------------------------------------------------------------------------
    def set_views(self, scenes: Sequence[Optional[BetaParams]], viewmats: Tensor, Ks: Tensor,
                  times: Optional[Sequence[float]] = None) -> None:
        """Slot k renders scenes[k] (None: empty) from viewmats[k] / Ks[k]
        (host or device tensors) at time times[k].  Queued on the current
        stream; the scenes' tensors must stay alive until the batch ran."""
        n = len(scenes)
        assert 1 <= n <= self.S, (n, self.S)
        ents = (_ViewScene * self.S)(*[scene_entry(scenes[k] if k < n else None) for k in range(self.S)])
        for k in range(n):
            if scenes[k] is not None:
                assert scenes[k].xyz.shape[0] <= self.R, (scenes[k].xyz.shape[0], self.R)
        ctypes.memmove(self._h_table.data_ptr(), ctypes.addressof(ents), ctypes.sizeof(ents))
        vm = viewmats.detach().to("cpu", torch.float32).reshape(n, 16)
        K = Ks.detach().to("cpu", torch.float32).reshape(n, 9)
        h = self._h_cams
        h.zero_()
        eye4, eye3 = torch.eye(4).reshape(16), torch.eye(3).reshape(9)
        h[:, :16] = eye4  # unused slots: a valid (identity) camera
        h[:, 16:25] = eye3
        h[:n, :16] = vm
        h[:n, 16:25] = K
        if times is not None:
            h[:n, 28] = torch.tensor(list(times), dtype=torch.float32)
        self.table.copy_(self._h_table, non_blocking=True)
        d = h.to(self.device, non_blocking=True)
        self.viewmats.copy_(d[:, :16].reshape(self.S, 4, 4))
        self.Ks.copy_(d[:, 16:25].reshape(self.S, 3, 3))
        self.times.copy_(d[:, 28])
        # the camera centres by the same device op as render_views / render_batch
        self.cam_pos.copy_(camera_centers(self.viewmats))
        self.views = list(scenes)

    # -- the launch sequence ---------------------------------------------------------------
    def run(self) -> None:
        """Enqueue the batch on the current stream (no host synchronisation)."""
        L = lib()
        dev, S, R, cap = self.device, self.S, self.R, self.cap
        st = stream_of(self.images)
        radii = torch.empty((S, R), dtype=torch.int32, device=dev)
        means2d = torch.empty((S, R, 2), dtype=torch.float32, device=dev)
        depths = torch.empty((S, R), dtype=torch.float32, device=dev)
        conics = torch.empty((S, R, 3), dtype=torch.float32, device=dev)
        opac = torch.empty((S, R), dtype=torch.float32, device=dev)
        betas = torch.empty((S, R), dtype=torch.float32, device=dev)
        colors = torch.empty((S, R, 3), dtype=torch.float32, device=dev)
        _call("ubs_beta_project_fwd_table", S, R, ptr(self.table), ptr(self.cam_pos), ptr(self.times),
              ptr(self.viewmats), ptr(self.Ks), ptr(self.wh), self.eps2d, self.near, self.far, self.rclip,
              ptr(radii), ptr(means2d), ptr(depths), ptr(conics), ptr(opac), ptr(betas), ptr(colors), st)
        # tile counts, their total (device) and the depth order of the (view, primitive) pairs
        tpg = torch.empty((S, R), dtype=torch.int32, device=dev)
        cum = torch.empty((S * R,), dtype=torch.int64, device=dev)
        self.info.zero_()
        ws_bytes = int(L.ubs_isect_count_workspace_bytes(S, R))
        ws = torch.empty((max(ws_bytes, 1),), dtype=torch.uint8, device=dev)
        _call("ubs_isect_tiles_count_range", S, R, ptr(means2d), ptr(radii), ptr(depths), self.ts, self.tw,
              self.th, ptr(tpg), ptr(cum), ptr(self.info), None, ptr(ws), ws_bytes, st)
        dkeys = torch.empty((S * R,), dtype=torch.int64, device=dev)
        dvals = torch.empty((S * R,), dtype=torch.int32, device=dev)
        _call("ubs_isect_depth_keys", S, R, ptr(radii), ptr(depths), ptr(dkeys), ptr(dvals), st)
        order = self._sort(dkeys, dvals, S * R, None, 0, 32)[1]
        ows = int(L.ubs_isect_ordered_workspace_bytes(S, R))
        ow = torch.empty((max(ows, 1),), dtype=torch.uint8, device=dev)
        _call("ubs_isect_scan_ordered", S, R, ptr(tpg), ptr(order), ptr(cum), ptr(ow), ows, st)
        # entries up to the capacity, sorted on the camera/tile bits, tile ranges
        ids = torch.empty((cap,), dtype=torch.int64, device=dev)
        flat = torch.empty((cap,), dtype=torch.int32, device=dev)
        _call("ubs_isect_tiles_emit_ordered_cap", S, R, ptr(means2d), ptr(radii), ptr(depths), ptr(order), ptr(cum),
              self.ts, self.tw, self.th, ptr(ids), ptr(flat), cap, st)
        ids, flat = self._sort(ids, flat, cap, self.info, 32, self.end_bit)
        offsets = torch.empty((S, self.th, self.tw), dtype=torch.int32, device=dev)
        _call("ubs_isect_offset_encode_dev", cap, ptr(self.info), ptr(ids), S, self.tw, self.th, ptr(offsets), st)
        bws = blend_workspace_bytes(S, R, self.tw, self.th, False)
        bw = torch.empty((bws,), dtype=torch.uint8, device=dev)
        last = torch.empty((S, self.H, self.W), dtype=torch.int32, device=dev)
        _call("ubs_rasterize_to_pixels_fwd_dev", S, R, cap, ptr(self.info), 3, ptr(means2d), ptr(conics),
              ptr(colors), ptr(opac), ptr(betas), None, None, self.W, self.H, self.ts, self.tw, self.th,
              ptr(offsets), ptr(flat), ptr(self.images), ptr(self.alphas), ptr(last), None, ptr(bw), bws, st)

    @staticmethod
    def _sort(keys: Tensor, vals: Tensor, n_cap: int, n_dev: Optional[Tensor], begin_bit: int, end_bit: int):
        L = lib()
        st = stream_of(keys)
        k_alt, v_alt = torch.empty_like(keys), torch.empty_like(vals)
        wsb = int(L.ubs_radix_sort_workspace_bytes(n_cap, begin_bit, end_bit))
        ws = torch.empty((max(wsb, 1),), dtype=torch.uint8, device=keys.device)
        sel = ctypes.c_int(0)
        if n_dev is None:
            _call("ubs_radix_sort_pairs", n_cap, ptr(keys), ptr(k_alt), ptr(vals), ptr(v_alt), begin_bit, end_bit,
                  ptr(ws), wsb, ctypes.byref(sel), st)
        else:
            _call("ubs_radix_sort_pairs_dev", n_cap, ptr(n_dev), ptr(keys), ptr(k_alt), ptr(vals), ptr(v_alt),
                  begin_bit, end_bit, ptr(ws), wsb, ctypes.byref(sel), st)
        return (k_alt, v_alt) if sel.value else (keys, vals)

    # -- graph -----------------------------------------------------------------------------
    def capture(self, pool=None) -> None:
        """Record run() as a graph (after one eager run on the same inputs:
        the first launches load code objects, which a capture cannot)."""
        self.run()
        torch.cuda.current_stream(self.device).synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, pool=pool):  # captured on a side stream, replayed on the current one
            self.run()
        self.graph = g

    def replay(self) -> None:
        assert self.graph is not None, "capture() first"
        self.graph.replay()

    def n_isects(self) -> int:
        """The last batch's intersection count (a host read: call it after
        the batch finished)."""
        return int(self.info[0].item())

    def overflowed(self) -> bool:
        return self.n_isects() > self.cap
