"""Graph-replayed serving batches (gsplat.batch_graph, RenderServer with
graphs=True): fixed-shape batches over several resident scenes with the
intersection lists sized on the device.

* eager and graph-replayed batches give every view's image bit for bit as
  render_views renders that view alone, over replays with other cameras and
  scene assignments, eager work and host synchronisations in between;
* empty slots render the background; a capacity that is too small is
  reported as an overflow, and the server then renders the batch again
  through the host-sized path (same images);
* at full size, 4 resident 2M-primitive scenes (two 6-D, two 7-D) under the
  device's real free-memory budget: every served image equals the view
  rendered alone.
"""
import numpy as np
import pytest
import torch

from gsplat import scenes
from gsplat.batch_graph import ViewBatchGraph
from gsplat.beta import BetaParams, render_views
from gsplat.serve import RenderServer, SceneStore, ViewRequest

pytestmark = pytest.mark.gpu


def _alone(p, vm, K, W, H, t, device):
    ts = torch.tensor([t], dtype=torch.float32, device=device) if p.D == 7 else None
    with torch.no_grad():
        img, alpha, _ = render_views(p, vm.to(device)[None].contiguous(), K.to(device)[None].contiguous(), W, H,
                                     timestamps=ts)
    return img[0], alpha[0]


def _cams(n, W, H, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    vms, Ks, ts = [], [], []
    for _ in range(n):
        d = rng.normal(0, 1, 3)
        vms.append(torch.tensor(scenes.look_at_viewmat(rng.uniform(3, 5) * d / np.linalg.norm(d)),
                                dtype=torch.float32))
        Ks.append(torch.tensor(scenes.intrinsics(W, H), dtype=torch.float32))
        ts.append(float(rng.uniform(0, 1)))
    return torch.stack(vms), torch.stack(Ks), ts


@pytest.fixture(scope="module")
def two_scenes(device):
    a = BetaParams.from_scene(scenes.make_scene(20_000, 6, seed=31), device, requires_grad=False)
    b = BetaParams.from_scene(scenes.make_scene(15_000, 7, seed=32), device, requires_grad=False, use_quats=False)
    return a, b


def test_graph_batch_matches_views_alone(device, two_scenes):
    a, b = two_scenes
    W, H = 200, 150
    g = ViewBatchGraph(device, 4, W, H, row_stride=20_000, cap=2_000_000)
    rounds = [([a, b, a], 1), ([b, b, a, None], 2), ([a], 3), ([b, a, a, b], 4)]
    for k, (assign, seed) in enumerate(rounds):
        vms, Ks, ts = _cams(len(assign), W, H, seed)
        print(f"round {k}: slots {[None if p is None else p.D for p in assign]}", flush=True)
        g.set_views(assign, vms, Ks, ts)
        if k == 0:
            g.run()  # eager
        elif g.graph is None:
            g.capture()
            g.replay()
        else:
            g.replay()
        torch.cuda.synchronize()
        assert not g.overflowed() and g.n_isects() > 0
        for s, p in enumerate(assign):
            if p is None:
                assert float(g.images[s].abs().max()) == 0.0 and float(g.alphas[s].abs().max()) == 0.0
                continue
            img, alpha = _alone(p, vms[s], Ks[s], W, H, ts[s], device)
            assert torch.equal(g.images[s], img), (k, s)
            assert torch.equal(g.alphas[s], alpha), (k, s)
        # eager work and a host read between replays
        _ = (torch.rand(1000, device=device) * 2).sum().item()


def test_graph_batch_overflow_is_reported(device, two_scenes):
    a, b = two_scenes
    W, H = 128, 96
    g = ViewBatchGraph(device, 2, W, H, row_stride=20_000, cap=1000)
    vms, Ks, ts = _cams(2, W, H, 7)
    g.set_views([a, b], vms, Ks, ts)
    g.run()
    torch.cuda.synchronize()
    assert g.overflowed() and g.n_isects() > 1000


def _requests(n, names, sizes, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    reqs = []
    for i in range(n):
        W, H = sizes[int(rng.integers(0, len(sizes)))]
        d = rng.normal(0, 1, 3)
        vm = scenes.look_at_viewmat(rng.uniform(3, 5) * d / np.linalg.norm(d))
        reqs.append(ViewRequest(names[int(rng.integers(0, len(names)))], torch.tensor(vm, dtype=torch.float32),
                                torch.tensor(scenes.intrinsics(W, H), dtype=torch.float32), W, H,
                                t=float(rng.uniform(0, 1)), rid=i))
    return reqs


def _serve(store, reqs, **kw):
    srv = RenderServer(store, graphs=True, **kw)
    for r in reqs:
        srv.submit(r)
    out = {r.rid: r for r in srv.drain()}
    assert len(out) == len(reqs)
    return out, srv


def test_server_graph_batches_cross_scene(device, two_scenes):
    a, b = two_scenes
    st = SceneStore(device)
    st.add("s6", a)
    st.add("s7", b)
    vm = torch.tensor(scenes.look_at_viewmat(np.array([0.0, 0.5, 4.0])), dtype=torch.float32)
    for name in st.scenes:
        st.calibrate(name, vm, torch.tensor(scenes.intrinsics(1920, 1080), dtype=torch.float32))
    reqs = _requests(14, ["s6", "s7"], [(128, 96), (200, 150)], 5)
    out, srv = _serve(st, reqs, max_views=8, streams=2)
    assert srv.overflows == 0
    assert srv.batches < len(reqs)  # views of both scenes share batches
    mixed = [r for r in out.values() if r.batch_views > 1]
    assert mixed
    for r in reqs:
        img, alpha = _alone(st.scenes[r.scene], r.viewmat, r.K, r.width, r.height, r.t, device)
        assert torch.equal(out[r.rid].image, img), r.rid
        assert torch.equal(out[r.rid].alpha, alpha), r.rid


def test_server_graph_overflow_renders_again(device, two_scenes):
    a, b = two_scenes
    st = SceneStore(device)
    st.add("s6", a)
    st.add("s7", b)
    reqs = _requests(4, ["s6", "s7"], [(128, 96)], 6)
    out, srv = _serve(st, reqs, max_views=4, streams=1, cap_margin=1e-4)
    assert srv.overflows >= 1
    for r in reqs:
        img, _ = _alone(st.scenes[r.scene], r.viewmat, r.K, r.width, r.height, r.t, device)
        assert torch.equal(out[r.rid].image, img), r.rid


def test_full_size_four_resident_scenes_real_budget(device):
    """4 resident scenes of 2M primitives (two 6-D static, two 7-D dynamic),
    admission by the device's free memory, graph batches over all scenes at
    640x480 / 1280x720 / 1920x1080: every image equals the view alone."""
    st = SceneStore(device)
    for name, D, seed in (("s6a", 6, 0), ("s6b", 6, 1), ("s7a", 7, 2), ("s7b", 7, 3)):
        st.add(name, BetaParams.from_scene(scenes.make_scene(2_000_000, D, seed=seed), device, requires_grad=False))
        print(f"resident {name}", flush=True)
    vm = torch.tensor(scenes.look_at_viewmat(np.array([0.0, 0.5, 4.0])), dtype=torch.float32)
    for name in st.scenes:
        st.calibrate(name, vm, torch.tensor(scenes.intrinsics(1920, 1080), dtype=torch.float32))
    reqs = _requests(12, list(st.scenes), [(640, 480), (1280, 720), (1920, 1080)], 8)
    out, srv = _serve(st, reqs, max_views=8, streams=2)
    print(f"batches {srv.batches}, overflows {srv.overflows}, resident {st.resident_bytes() / 2**30:.2f} GiB",
          flush=True)
    assert st.resident_bytes() > 4 * 2_000_000 * 140
    assert srv.batches < len(reqs)
    for r in reqs:
        img, _ = _alone(st.scenes[r.scene], r.viewmat, r.K, r.width, r.height, r.t, device)
        assert torch.equal(out[r.rid].image, img), r.rid
