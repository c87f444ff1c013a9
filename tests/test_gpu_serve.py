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
from gsplat.cuda._wrapper import raster_fwd_raw
from gsplat.serve import RenderServer, SceneStore, ViewRequest, latency_summary

from util import BRANCH_FLIP_FRAC, check_branch_pixels

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


def test_full_size_mix_matches_alone_and_fp64(device):
    """Serving at full size: two resident 2M-primitive scenes (6-D and 7-D),
    requests at 640x480 / 1280x720 / 1920x1080: every image is bit-identical
    alone and in the mixed batches, and a 1080p view matches the FP64 oracle
    to 1e-4 max-abs on 64 random tiles (pixels away from the T cut-off),
    blending the served view's own tile lists."""
    from oracle import ubs_oracle as O

    st = SceneStore(device)
    scs = {}
    for name, D, seed in (("big6", 6, 40), ("big7", 7, 41)):
        scs[name] = scenes.make_scene(2_000_000, D, seed=seed)
        st.add(name, BetaParams.from_scene(scs[name], device, requires_grad=False, use_quats=True))
    rng = np.random.Generator(np.random.PCG64(9))
    sizes = [(640, 480), (1280, 720), (1920, 1080)]
    reqs = []
    for i in range(6):
        W, H = sizes[i % 3]
        d = rng.normal(0, 1, 3)
        vm = scenes.look_at_viewmat(rng.uniform(3, 5) * d / np.linalg.norm(d))
        reqs.append(ViewRequest(["big6", "big7"][i % 2], torch.tensor(vm, dtype=torch.float32),
                                torch.tensor(scenes.intrinsics(W, H), dtype=torch.float32), W, H,
                                t=float(rng.uniform(0, 1)), rid=i))
    print("full-size serving: 2 x 2M primitives, 6 requests", flush=True)
    mixed, srv = _serve(st, reqs, max_views=16, streams=2)
    alone, _ = _serve(st, reqs, max_views=1, streams=1)
    assert srv.batches == 2  # one batch per scene
    for r in reqs:
        assert torch.equal(mixed[r.rid].image, alone[r.rid].image), r.rid
    # FP64 oracle on 64 tiles of the 1080p 7-D view (rid 5), in the served view's tile lists
    r = reqs[5]
    p = st.scenes[r.scene]
    ts = torch.tensor([r.t], dtype=torch.float32, device=device)
    with torch.no_grad():
        img, _, meta = render_views(p, r.viewmat.to(device)[None].contiguous(), r.K.to(device)[None].contiguous(),
                                    r.width, r.height, timestamps=ts)
    assert torch.equal(img[0], mixed[r.rid].image)
    sc = scs[r.scene]
    tw, th = -(-r.width // 16), -(-r.height // 16)
    tm = np.zeros(th * tw, bool)
    tm[np.random.Generator(np.random.PCG64(1)).choice(th * tw, 64, replace=False)] = True
    tm = tm.reshape(th, tw)
    struct = tuple(meta[k].cpu().numpy() for k in ("tiles_per_gauss", "isect_ids", "flatten_ids", "isect_offsets"))
    leaves = {k: getattr(sc, k) for k in ("xyz", "mean_rest", "scale_raw", "l_triangle", "opacity_raw", "beta_raw",
                                         "rgb", "quats")}
    ref, _, rm = O.beta_step(leaves, r.viewmat.numpy(), r.K.numpy(), r.width, r.height,
                             np.zeros((r.height, r.width, 3)), timestamp=r.t, dtype=np.float64, structure=struct,
                             want_grads=False, tile_mask=tm)
    _, _, _, mT, mS = O.raster_fwd(rm["means2d"], rm["conics"], rm["colors"], rm["opacities"], rm["betas"], r.width,
                                   r.height, 16, rm["isect_offsets"], rm["flatten_ids"], masks=tm[None], margins=True)
    pm = np.kron(tm, np.ones((16, 16), bool))[:r.height, :r.width]
    safe = pm & (mT[0] > 1e-3) & (mS[0] > 1e-6)
    assert safe.sum() > 0.9 * pm.sum()
    err = np.abs(mixed[r.rid].image.cpu().numpy() - ref)[safe].max()
    assert err <= 1e-4, err
    # the pixels near a branch point: the float32 oracle's blend of the served
    # view's own projected lists, colour and last id
    g32 = {k: meta[k].cpu().numpy() for k in ("means2d", "conics", "opacities", "betas", "isect_offsets",
                                                "flatten_ids")}
    rgb = p.rgb[None].contiguous()
    oc, _, olast = O.raster_fwd(g32["means2d"], g32["conics"], rgb.cpu().numpy(), g32["opacities"], g32["betas"],
                                r.width, r.height, 16, g32["isect_offsets"], g32["flatten_ids"], masks=tm[None])
    with torch.no_grad():
        _, _, last, _ = raster_fwd_raw(meta["means2d"], meta["conics"], rgb, meta["opacities"], meta["betas"], None,
                                       None, r.width, r.height, 16, meta["isect_offsets"], meta["flatten_ids"])
    check_branch_pixels(mixed[r.rid].image.cpu().numpy(), last[0].cpu().numpy(), oc[0], olast[0], pm, safe,
                        atol=5e-5, what="served 1080p view near branch points")
    far64 = pm & (np.abs(mixed[r.rid].image.cpu().numpy() - ref).max(axis=-1) > 1e-4)
    assert far64.sum() <= max(1, int(BRANCH_FLIP_FRAC * pm.sum())), int(far64.sum())
