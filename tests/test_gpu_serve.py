"""GPU checks of the serving loop (gsplat.serve): every view's image is the
same, bit for bit, alone and in a mixed batch (two resident scenes, 6-D and
7-D, mixed resolutions, two streams), and equals a single-view render_views
of the same camera; a memory budget that forces one view per batch changes
nothing."""
import numpy as np
import pytest
import torch

from gsplat import scenes
from gsplat.beta import BetaParams, render_views
from gsplat.serve import RenderServer, SceneStore, ViewRequest, latency_summary

pytestmark = pytest.mark.gpu

SIZES = [(128, 96), (200, 150), (64, 64)]


def _requests(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    reqs = []
    for i in range(n):
        scene = ["s6", "s7"][int(rng.integers(0, 2))]
        W, H = SIZES[int(rng.integers(0, 3))]
        d = rng.normal(0, 1, 3)
        vm = scenes.look_at_viewmat(rng.uniform(3, 5) * d / np.linalg.norm(d))
        reqs.append(ViewRequest(scene, torch.tensor(vm, dtype=torch.float32), torch.tensor(
            scenes.intrinsics(W, H), dtype=torch.float32), W, H, t=float(rng.uniform(0, 1)), rid=i))
    return reqs


@pytest.fixture(scope="module")
def store(device):
    st = SceneStore(device)
    st.add("s6", BetaParams.from_scene(scenes.make_scene(20_000, 6, seed=21), device, requires_grad=False))
    st.add("s7", BetaParams.from_scene(scenes.make_scene(15_000, 7, seed=22), device, requires_grad=False,
                                       use_quats=False))
    return st


def _serve(store, reqs, **kw):
    srv = RenderServer(store, **kw)
    for r in reqs:
        srv.submit(ViewRequest(r.scene, r.viewmat, r.K, r.width, r.height, r.t, rid=r.rid))
    res = srv.drain()
    torch.cuda.synchronize()
    return {r.rid: r for r in res}, srv


def test_results_same_alone_and_in_the_mix(store):
    reqs = _requests(14, 0)
    mixed, srv = _serve(store, reqs, max_views=8, streams=2)
    alone, _ = _serve(store, reqs, max_views=1, streams=1)
    assert srv.batches < len(reqs)  # views were batched
    assert max(r.batch_views for r in mixed.values()) > 1
    for r in reqs:
        a, b = mixed[r.rid], alone[r.rid]
        assert a.image.shape == (r.height, r.width, 3)
        assert torch.equal(a.image, b.image) and torch.equal(a.alpha, b.alpha), r.rid
        p = store.scenes[r.scene]
        ts = torch.tensor([r.t], dtype=torch.float32, device=store.device) if p.D == 7 else None
        with torch.no_grad():
            img, alpha, _ = render_views(p, r.viewmat.to(store.device)[None].contiguous(),
                                         r.K.to(store.device)[None].contiguous(), r.width, r.height, timestamps=ts)
        assert torch.equal(a.image, img[0]) and torch.equal(a.alpha, alpha[0]), r.rid
        assert float(a.alpha.max()) > 0  # something was drawn
    s = latency_summary(list(mixed.values()), 1.0)
    assert s["views"] == len(reqs)


def test_tight_memory_budget_serves_one_view_at_a_time(store):
    reqs = _requests(6, 1)
    tight, srv = _serve(store, reqs, max_views=8, streams=2, free_bytes_fn=lambda: 0)
    loose, _ = _serve(store, reqs, max_views=8, streams=1)
    assert all(r.batch_views == 1 for r in tight.values())
    for r in reqs:
        assert torch.equal(tight[r.rid].image, loose[r.rid].image)


def test_calibrate_measures_isects(store):
    vm = torch.tensor(scenes.look_at_viewmat(np.array([0.0, 0.5, 4.0])), dtype=torch.float32)
    k = store.calibrate("s6", vm, torch.tensor(scenes.intrinsics(1920, 1080), dtype=torch.float32))
    assert 1.0 <= k < 200.0 and store.isects_per_prim["s6"] == k
