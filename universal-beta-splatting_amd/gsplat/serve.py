"""Render serving on one GPU: Beta scenes resident in HBM, view requests of
mixed resolution admitted by free memory and batched continuously.

The reference serves views one at a time (BetaModel.view through the
viewer, scene/beta_model.py:724-804, and Scene.eval's per-camera loop,
scene/__init__.py:170-190).  Here every scene stays resident as its leaf
tensors (``SceneStore``), and ``RenderServer`` turns the pending requests
into batches: all pending views of one scene go through one fused Beta
projection (per-view query, time and image size), then one intersection /
sort / blend pass per resolution group (``beta.render_batch``).  Batches
run on a small pool of HIP streams (default priority), so a batch formed
from requests that arrived while another was on the GPU overlaps with it.
A request is admitted when its estimated working memory fits in the free
HBM (device free memory plus the allocator's unused cache, minus a
reserve); the estimate per view follows the scene's observed
intersections per primitive.

Every view's image depends only on its own camera, time and scene, so a
result is the same alone and in any batch (the projection, the per-view
tile lists and the blend of a view read nothing of the other views).
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass
from typing import Callable, Deque, Dict, List, Optional, Tuple

import torch
from torch import Tensor

from .beta import BetaParams, render_batch


# Capacity bound of a graph batch's intersection lists (the sort's key limit).
GRAPH_MAX_ISECTS = (1 << 30) - 1


@dataclass
class ViewRequest:
    """One view of a resident scene: world->camera [4,4] row-major,
    intrinsics [3,3], image size and (7-D scenes) time."""
    scene: str
    viewmat: Tensor
    K: Tensor
    width: int
    height: int
    t: float = 0.0
    arrival: float = 0.0   # host clock (time.perf_counter) when the request arrived
    rid: int = -1


@dataclass
class ViewResult:
    rid: int
    scene: str
    image: Optional[Tensor]  # [H,W,3] (None when the server keeps no images)
    alpha: Optional[Tensor]  # [H,W,1]
    arrival: float
    done: float              # host clock when the batch's GPU work was seen complete
    batch_views: int
    pixels: int = 0

    @property
    def latency_ms(self) -> float:
        return (self.done - self.arrival) * 1e3


class SceneStore:
    """Scenes resident on one device as leaf tensors (no gradients)."""

    def __init__(self, device):
        self.device = torch.device(device)
        self.scenes: Dict[str, BetaParams] = {}
        self.isects_per_prim: Dict[str, float] = {}

    def add(self, name: str, params: BetaParams) -> None:
        leaves = {k: v.detach().to(self.device, torch.float32).contiguous() for k, v in params.leaves().items()}
        self.scenes[name] = BetaParams(**leaves)
        self.isects_per_prim[name] = 16.0  # a prior; calibrate() measures it

    @torch.no_grad()
    def calibrate(self, name: str, viewmat: Tensor, K: Tensor, width: int = 1920, height: int = 1080,
                  t: float = 0.0) -> float:
        """Render one view and keep its intersections per primitive
        (scaled to a 1080p view) as the scene's memory estimate."""
        from .beta import render_views
        p = self.scenes[name]
        ts = torch.tensor([t], dtype=torch.float32, device=self.device) if p.D == 7 else None
        _, _, meta = render_views(p, viewmat.to(self.device, torch.float32)[None].contiguous(),
                                  K.to(self.device, torch.float32)[None].contiguous(), width, height, timestamps=ts)
        area = (width * height) / (1920.0 * 1080.0)
        self.isects_per_prim[name] = max(1.0, meta["isect_ids"].numel() / max(1, p.xyz.shape[0]) / area)
        return self.isects_per_prim[name]

    def remove(self, name: str) -> None:
        """Drop a scene.  A batch of it still in flight on a RenderServer keeps
        its tensors alive until the server collects that batch."""
        self.scenes.pop(name, None)
        self.isects_per_prim.pop(name, None)

    def resident_bytes(self) -> int:
        return sum(t.numel() * t.element_size() for p in self.scenes.values() for t in p.leaves().values())

    def view_isects(self, name: str, width: int, height: int) -> float:
        """Estimated intersections of one view of `name` at width x height.
        A primitive's footprint in pixels scales with the image width (the
        focal length); with k intersections per primitive measured at 1080p
        (calibrate), its tile box is about sqrt(k) tiles wide there, so
        (s (sqrt(k) - 1) + 1)^2 tiles at scale s = width / 1920."""
        k = max(1.0, self.isects_per_prim[name])
        sc = width / 1920.0
        return self.scenes[name].xyz.shape[0] * (sc * (math.sqrt(k) - 1.0) + 1.0) ** 2

    def view_bytes(self, name: str, width: int, height: int) -> int:
        """Working memory one view of `name` needs: projected rows (36 B per
        primitive), the intersection keys/values in both sort buffers (24 B
        each, view_isects of them), count/prefix arrays, and the image."""
        N = self.scenes[name].xyz.shape[0]
        return int(36 * N + 12 * N + 24 * self.view_isects(name, width, height) + 24 * width * height)


def free_hbm_bytes(device) -> int:
    free, _ = torch.cuda.mem_get_info(device)
    return int(free + torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device))


@dataclass
class _InFlight:
    stream: torch.cuda.Stream
    event: torch.cuda.Event
    reqs: List[ViewRequest]
    images: List[Tensor]
    alphas: List[Tensor]
    bytes: int
    params: List[BetaParams]  # held until the batch is collected: a scene removed meanwhile stays allocated
    graph: Optional[object] = None  # the ViewBatchGraph that rendered it (its count is checked at collect)


class RenderServer:
    """Continuous batching of view requests over a SceneStore.

    ``submit`` queues a request; ``poll`` collects finished batches and
    launches new ones on free streams; ``drain`` polls until every
    submitted request has a result.  ``max_views`` bounds a batch,
    ``streams`` the batches in flight; with max_views=1 and streams=1 the
    server renders one view at a time, in arrival order (the reference's
    structure)."""

    def __init__(self, store: SceneStore, max_views: int = 16, streams: int = 2,
                 hbm_reserve_bytes: int = 1 << 30, tile_size: int = 16,
                 render_fn: Optional[Callable] = None, free_bytes_fn: Optional[Callable[[], int]] = None,
                 keep_images: bool = True, graphs: bool = False, cap_margin: float = 1.5):
        self.store = store
        # graphs: batches of one image size over any resident scenes run as a
        # replayed hipGraph per (size, slots) with device-sized lists
        # (gsplat.batch_graph); otherwise one scene's views per batch through
        # render_batch (host-sized lists, any mix of sizes)
        self.graphs = bool(graphs) and render_fn is None
        self.cap_margin = float(cap_margin)
        self._pipes: List[Dict[Tuple[int, int, int, int], object]] = []
        self._pools: List[object] = []
        self.overflows = 0
        self.keep_images = keep_images
        self.max_views = max(1, int(max_views))
        self.reserve = int(hbm_reserve_bytes)
        self.tile_size = tile_size
        self.render_fn = render_fn or render_batch
        self.free_bytes_fn = free_bytes_fn or (lambda: free_hbm_bytes(store.device))
        dev = store.device
        self.streams = [torch.cuda.Stream(device=dev) for _ in range(max(1, int(streams)))] \
            if dev.type == "cuda" else [None] * max(1, int(streams))
        self.pending: Deque[ViewRequest] = collections.deque()
        self.inflight: List[Optional[_InFlight]] = [None] * len(self.streams)
        self._pipes = [dict() for _ in self.streams]
        self._pools = [None] * len(self.streams)
        self.results: List[ViewResult] = []
        self._next_rid = 0
        self.batches = 0

    # -- requests ---------------------------------------------------------------------------
    def submit(self, req: ViewRequest) -> int:
        if req.scene not in self.store.scenes:
            raise KeyError(f"scene {req.scene!r} is not resident")
        if req.rid < 0:
            req.rid = self._next_rid
        self._next_rid = max(self._next_rid, req.rid) + 1
        if not req.arrival:
            req.arrival = time.perf_counter()
        self.pending.append(req)
        return req.rid

    def busy(self) -> bool:
        return bool(self.pending) or any(f is not None for f in self.inflight)

    # -- scheduling -------------------------------------------------------------------------
    def _form_batch(self, budget: int) -> Tuple[List[ViewRequest], int]:
        """The oldest pending request, then every pending request that can
        share its batch in arrival order -- with graphs, those of the same
        image size over any scene; without, those of the same scene -- while
        the batch fits max_views and the memory budget (the oldest request
        is always taken, so the queue cannot starve)."""
        head = self.pending[0]
        same = (lambda r: (r.width, r.height) == (head.width, head.height)) if self.graphs else \
            (lambda r: r.scene == head.scene)
        most = self._graph_max_views(head.width, head.height) if self.graphs else self.max_views
        batch, used, keep = [], 0, collections.deque()
        while self.pending:
            r = self.pending.popleft()
            need = self.store.view_bytes(r.scene, r.width, r.height)
            if same(r) and len(batch) < most and (not batch or used + need <= budget):
                batch.append(r)
                used += need
            else:
                keep.append(r)
        self.pending = keep
        return batch, used

    def _graph_view_isects(self, W: int, H: int) -> float:
        """A graph slot's share of the lists: the largest estimate over the
        resident scenes (any scene may take any slot)."""
        return max(self.store.view_isects(k, W, H) for k in self.store.scenes)

    def _graph_max_views(self, W: int, H: int) -> int:
        """Most views of one graph batch at W x H: max_views, lowered so that
        the lists' capacity stays within what the sort holds."""
        per = max(1.0, self.cap_margin * self._graph_view_isects(W, H))
        return max(1, min(self.max_views, int(GRAPH_MAX_ISECTS // per)))

    def prepare_graphs(self, sizes) -> int:
        """Capture, on every stream, the graph of each batch shape the server
        can form at these image sizes (every view count up to the size's
        limit), so that no capture falls inside serving.  Returns the number
        of graphs captured."""
        import numpy as np

        from . import scenes as _scenes

        if not self.graphs:
            return 0
        p = next(iter(self.store.scenes.values()))
        vm = torch.tensor(_scenes.look_at_viewmat(np.array([0.0, 0.4, 4.0])), dtype=torch.float32)
        made = 0
        for W, H in sizes:
            K = torch.tensor(_scenes.intrinsics(W, H), dtype=torch.float32)
            for n in range(1, self._graph_max_views(W, H) + 1):
                for slot, s in enumerate(self.streams):
                    g = self._pipe(slot, W, H, n)
                    if not g.needs_capture:
                        continue
                    with torch.no_grad(), torch.cuda.stream(s):
                        g.set_views([p] * n, vm[None].expand(n, 4, 4), K[None].expand(n, 3, 3), [0.5] * n)
                        if self._pools[slot] is None:
                            self._pools[slot] = torch.cuda.graph_pool_handle()
                        g.capture(pool=self._pools[slot])
                        g.needs_capture = False
                    made += 1
        torch.cuda.synchronize(self.store.device)
        return made

    def _pipe(self, slot: int, W: int, H: int, n_views: int):
        """The stream slot's graph for (W, H) and n_views slots, captured on
        first use (or by prepare_graphs); its lists hold cap_margin times the
        largest resident scene's estimated intersections per slot."""
        from .batch_graph import ViewBatchGraph

        S = int(n_views)
        st = self.store
        R = max(p.xyz.shape[0] for p in st.scenes.values())
        key = (W, H, S, R)
        g = self._pipes[slot].get(key)
        if g is None:
            cap = min(GRAPH_MAX_ISECTS, int(math.ceil(self.cap_margin * self._graph_view_isects(W, H) * S)) + 1024)
            g = ViewBatchGraph(st.device, S, W, H, R, cap, tile_size=self.tile_size)
            g.needs_capture = True
            self._pipes[slot][key] = g
        return g

    def _launch(self, slot: int, batch: List[ViewRequest], nbytes: int) -> None:
        if self.graphs:
            return self._launch_graph(slot, batch, nbytes)
        params = self.store.scenes[batch[0].scene]
        s = self.streams[slot]
        dev = self.store.device
        caller = torch.cuda.current_stream(dev) if s is not None else None
        ctx = torch.cuda.stream(s) if s is not None else _nullctx()
        with torch.no_grad(), ctx:
            if s is not None:  # inputs made on the caller's stream are ready before this batch reads them
                s.wait_stream(caller)
            vms = torch.stack([r.viewmat.to(dev, torch.float32) for r in batch]).contiguous()
            Ks = torch.stack([r.K.to(dev, torch.float32) for r in batch]).contiguous()
            ts = None
            if params.D == 7:
                ts = torch.tensor([r.t for r in batch], dtype=torch.float32, device=dev)
            images, alphas = self.render_fn(params, vms, Ks, [(r.width, r.height) for r in batch], timestamps=ts,
                                            tile_size=self.tile_size)
            ev = torch.cuda.Event() if s is not None else None
            if ev is not None:
                ev.record(s)
        self.inflight[slot] = _InFlight(s, ev, batch, images, alphas, nbytes, [params])
        self.batches += 1

    def _launch_graph(self, slot: int, batch: List[ViewRequest], nbytes: int) -> None:
        # slots ordered by scene: consecutive views of one scene share the
        # per-primitive state in the projection
        batch = sorted(batch, key=lambda r: r.scene)
        W, H = batch[0].width, batch[0].height
        dev = self.store.device
        s = self.streams[slot]
        caller = torch.cuda.current_stream(dev)
        scenes_ = [self.store.scenes[r.scene] for r in batch]
        with torch.no_grad(), torch.cuda.stream(s):
            s.wait_stream(caller)
            g = self._pipe(slot, W, H, len(batch))
            g.set_views(scenes_, torch.stack([r.viewmat for r in batch]), torch.stack([r.K for r in batch]),
                        [r.t for r in batch])
            if g.needs_capture:
                if self._pools[slot] is None:
                    self._pools[slot] = torch.cuda.graph_pool_handle()
                g.capture(pool=self._pools[slot])
                g.needs_capture = False
            g.replay()
            # the graph's outputs are overwritten by its next replay: copies
            images = [g.images[k].clone() for k in range(len(batch))]
            alphas = [g.alphas[k].clone() for k in range(len(batch))]
            ev = torch.cuda.Event()
            ev.record(s)
        self.inflight[slot] = _InFlight(s, ev, batch, images, alphas, nbytes, scenes_, g)
        self.batches += 1

    def _rerender(self, f: _InFlight) -> None:
        """A graph batch whose intersections exceeded its lists' capacity:
        render its views again through the host-sized path (one scene at a
        time) and retire the graph, so the next batch of its shape is
        captured with twice the room."""
        self.overflows += 1
        g = f.graph
        for key, v in list(self._pipes[self.streams.index(f.stream)].items()):
            if v is g:
                del self._pipes[self.streams.index(f.stream)][key]
        self.cap_margin *= 2.0
        dev = self.store.device
        with torch.no_grad():
            for k, r in enumerate(f.reqs):
                p = self.store.scenes.get(r.scene) or f.params[k]
                ts = torch.tensor([r.t], dtype=torch.float32, device=dev) if p.D == 7 else None
                imgs, alphas = self.render_fn(p, r.viewmat.to(dev, torch.float32)[None].contiguous(),
                                              r.K.to(dev, torch.float32)[None].contiguous(), [(r.width, r.height)],
                                              timestamps=ts, tile_size=self.tile_size)
                f.images[k], f.alphas[k] = imgs[0], alphas[0]

    def _collect(self, slot: int, now: float) -> None:
        f = self.inflight[slot]
        if f.graph is not None and f.graph.overflowed():
            self._rerender(f)
            now = time.perf_counter()
        consumer = torch.cuda.current_stream(self.store.device) if f.stream is not None else None
        for r, img, a in zip(f.reqs, f.images, f.alphas):
            keep = self.keep_images
            if keep and consumer is not None and consumer != f.stream:
                # the results were allocated on the batch's stream: when the caller
                # drops them, their memory returns to that stream's pool only after
                # what the caller's stream has queued by then has run
                img.record_stream(consumer)
                a.record_stream(consumer)
            self.results.append(ViewResult(r.rid, r.scene, img if keep else None, a if keep else None, r.arrival,
                                           now, len(f.reqs), r.width * r.height))
        self.inflight[slot] = None

    def poll(self) -> int:
        """Collect finished batches, then start batches on free streams.
        Returns the number of results collected."""
        n0 = len(self.results)
        for k, f in enumerate(self.inflight):
            if f is not None and (f.event is None or f.event.query()):
                self._collect(k, time.perf_counter())
        for k in range(len(self.streams)):
            if not self.pending:
                break
            if self.inflight[k] is not None:
                continue
            held = sum(f.bytes for f in self.inflight if f is not None)
            budget = self.free_bytes_fn() - self.reserve - held
            head = self.pending[0]
            if self.store.view_bytes(head.scene, head.width, head.height) > budget and held > 0:
                break  # wait for a batch in flight to free its memory
            batch, used = self._form_batch(budget)
            self._launch(k, batch, used)
        return len(self.results) - n0

    def drain(self, timeout_s: float = 600.0) -> List[ViewResult]:
        t_end = time.perf_counter() + timeout_s
        while self.busy():
            if time.perf_counter() > t_end:
                raise TimeoutError("render server did not drain")
            self.poll()
        out, self.results = self.results, []
        return out


class _nullctx:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


def latency_summary(results: List[ViewResult], wall_s: float) -> dict:
    """p50 / p99 latency (ms), frames/s and MPix/s of a served run."""
    lat = sorted(r.latency_ms for r in results)
    if not lat:
        return {"views": 0}
    pix = sum(r.pixels for r in results)

    def q(p):
        return lat[min(len(lat) - 1, max(0, int(math.ceil(p * len(lat))) - 1))]
    return {"views": len(lat), "p50_ms": round(q(0.50), 3), "p99_ms": round(q(0.99), 3),
            "max_ms": round(lat[-1], 3), "frames_per_s": round(len(lat) / wall_s, 2),
            "mpix_per_s": round(pix / wall_s / 1e6, 2),
            "mean_batch_views": round(sum(r.batch_views for r in results) / len(results), 2)}
