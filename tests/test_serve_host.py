"""Scheduler logic of gsplat.serve on the CPU, with a stand-in render
function (no kernels): batches take the oldest request's scene and every
pending request of it in arrival order, bounded by max_views and the memory
budget; every request gets exactly one result, its own."""
import numpy as np
import pytest
import torch

from gsplat import scenes
from gsplat.beta import BetaParams
from gsplat.serve import RenderServer, SceneStore, ViewRequest, latency_summary


def _store():
    st = SceneStore("cpu")
    for name, D, n in (("a", 6, 100), ("b", 7, 50)):
        st.add(name, BetaParams.from_scene(scenes.make_scene(n, D, seed=n), "cpu", requires_grad=False,
                                           use_quats=False))
    return st


def _fake_render(calls):
    def render(params, vms, Ks, sizes, timestamps=None, tile_size=16):
        calls.append((params.D, len(sizes), None if timestamps is None else timestamps.tolist()))
        imgs = [torch.full((h, w, 3), float(vms[i, 0, 3])) for i, (w, h) in enumerate(sizes)]
        return imgs, [torch.ones((h, w, 1)) for (w, h) in sizes]
    return render


def _req(scene, rid, w=32, h=24, t=0.0):
    vm = torch.eye(4)
    vm[0, 3] = float(rid)  # tags the request's image
    return ViewRequest(scene, vm, torch.eye(3), w, h, t=t, arrival=1.0 + rid, rid=rid)


def test_batches_group_by_scene_in_arrival_order():
    calls = []
    srv = RenderServer(_store(), max_views=3, streams=1, hbm_reserve_bytes=0, render_fn=_fake_render(calls),
                       free_bytes_fn=lambda: 1 << 40)
    order = ["a", "b", "a", "a", "b", "a", "a"]
    for i, sc in enumerate(order):
        srv.submit(_req(sc, i, t=0.1 * i))
    res = srv.drain()
    assert sorted(r.rid for r in res) == list(range(len(order)))
    # batch 1: a0 a2 a3 (max 3); batch 2: b1 b4; batch 3: a5 a6
    assert [c[1] for c in calls] == [3, 2, 2]
    assert calls[1][0] == 7 and calls[1][2] == pytest.approx([0.1, 0.4])  # 7-D scene gets its times
    assert calls[0][2] is None
    for r in res:
        assert float(r.image[0, 0, 0]) == r.rid and r.pixels == 32 * 24
    assert {r.rid: r.batch_views for r in res} == {0: 3, 2: 3, 3: 3, 1: 2, 4: 2, 5: 2, 6: 2}


def test_memory_budget_limits_batches_but_never_starves():
    calls = []
    st = _store()
    need = st.view_bytes("a", 32, 24)
    srv = RenderServer(st, max_views=16, streams=1, hbm_reserve_bytes=0, render_fn=_fake_render(calls),
                       free_bytes_fn=lambda: int(2.5 * need))
    for i in range(5):
        srv.submit(_req("a", i))
    res = srv.drain()
    assert [c[1] for c in calls] == [2, 2, 1]
    assert sorted(r.rid for r in res) == list(range(5))
    calls.clear()
    srv = RenderServer(st, max_views=16, streams=1, hbm_reserve_bytes=0, render_fn=_fake_render(calls),
                       free_bytes_fn=lambda: 0)  # nothing fits: one view at a time, still served
    for i in range(3):
        srv.submit(_req("a", i))
    assert len(srv.drain()) == 3 and [c[1] for c in calls] == [1, 1, 1]


def test_per_view_server_is_fifo():
    calls = []
    srv = RenderServer(_store(), max_views=1, streams=1, render_fn=_fake_render(calls),
                       free_bytes_fn=lambda: 1 << 40)
    for i, sc in enumerate(["b", "a", "b"]):
        srv.submit(_req(sc, i))
    res = srv.drain()
    assert [r.rid for r in res] == [0, 1, 2] and [c[0] for c in calls] == [7, 6, 7]


def test_unknown_scene_and_view_bytes():
    st = _store()
    srv = RenderServer(st, render_fn=_fake_render([]), free_bytes_fn=lambda: 1 << 40)
    with pytest.raises(KeyError):
        srv.submit(_req("zzz", 0))
    assert st.view_bytes("a", 1920, 1080) > st.view_bytes("a", 640, 480) > 0
    assert st.resident_bytes() == sum(t.numel() * 4 for p in st.scenes.values() for t in p.leaves().values())


def test_latency_summary():
    from gsplat.serve import ViewResult
    res = [ViewResult(i, "a", None, None, 0.0, (i + 1) * 1e-3, 2, 100) for i in range(100)]
    s = latency_summary(res, wall_s=0.5)
    assert s["p50_ms"] == pytest.approx(50.0) and s["p99_ms"] == pytest.approx(99.0)
    assert s["frames_per_s"] == 200.0 and s["mpix_per_s"] == pytest.approx(100 * 100 / 0.5 / 1e6)
