"""Parity of the batched multi-view step at its full size: a 3M-primitive
6-D scene and 8 views cycling through 960x540 / 1280x720 / 1920x1080 (the
batched-training workload), one render_batch forward + backward of the
summed per-view L1 loss against the same views rendered one at a time, the
reference's per-view loop (train.py:111-128):

* every view's image and alpha are bit-identical to rendering that view
  alone (render_views with one camera): the fused projection, the tile
  lists and the blend of a view read nothing of the other views;
* every leaf gradient equals the sum of the 8 single-view gradients to a
  relative L2 of 1e-5 and a max-abs of 1e-4 of the largest entry: the
  blend's float atomics add the contributions in another order, so the
  sums agree to rounding, not bit for bit.
"""
import numpy as np
import pytest
import torch

from gsplat import beta, scenes

from util import rel_l2

pytestmark = pytest.mark.gpu

N, D, VIEWS = 3_000_000, 6, 8
SIZES = ((960, 540), (1280, 720), (1920, 1080))


def test_full_size_batch_matches_per_view_loop(device):
    sizes = [SIZES[c % len(SIZES)] for c in range(VIEWS)]
    sc = scenes.make_scene(N, D, seed=0)
    p = beta.BetaParams.from_scene(sc, device)
    vms, _, _ = scenes.make_cameras(VIEWS, sizes[0][0], sizes[0][1], seed=0)
    Ks = np.stack([scenes.intrinsics(w, h) for w, h in sizes]).astype(np.float32)
    vms_t, Ks_t = torch.tensor(vms, device=device), torch.tensor(Ks, device=device)
    g = np.random.Generator(np.random.PCG64(5))
    tgts = [torch.tensor(g.uniform(0, 1, (h, w, 3)), dtype=torch.float32, device=device) for w, h in sizes]
    names = [k for k, v in p.leaves().items() if v is not None]
    leaves = [p.leaves()[k] for k in names]

    imgs, alphas = beta.render_batch(p, vms_t, Ks_t, sizes)
    loss = sum((im - t).abs().mean() for im, t in zip(imgs, tgts))
    g_batch = torch.autograd.grad(loss, leaves)
    imgs = [im.detach() for im in imgs]
    alphas = [a.detach() for a in alphas]

    g_sum = None
    for c, (w, h) in enumerate(sizes):
        one, a1, _ = beta.render_views(p, vms_t[c:c + 1], Ks_t[c:c + 1], w, h)
        assert torch.equal(imgs[c], one[0].detach()), f"view {c} image"
        assert torch.equal(alphas[c], a1[0].detach()), f"view {c} alpha"
        gc = torch.autograd.grad((one[0] - tgts[c]).abs().mean(), leaves)
        g_sum = list(gc) if g_sum is None else [x + y for x, y in zip(g_sum, gc)]
        del one, a1, gc

    for name, a, b in zip(names, g_batch, g_sum):
        assert torch.isfinite(a).all(), name
        scale = b.abs().max().item()
        assert scale > 0, name
        rel = ((a.double() - b.double()).norm() / b.double().norm()).item()
        assert rel <= 1e-5, (name, rel)
        assert (a - b).abs().max().item() <= 1e-4 * scale, (name, (a - b).abs().max().item(), scale)


def _view_subset(meta, tm, N):
    """The primitives the marked tiles of one view blend, and that view's
    tile lists with flatten ids renumbered into the subset (entries of
    unmarked tiles, which a tile-masked blend never reads, point at 0)."""
    offs = meta["isect_offsets"].cpu().numpy()[0]
    flat = meta["flatten_ids"].cpu().numpy()
    starts = offs.ravel().astype(np.int64)
    ends = np.append(starts[1:], flat.size)
    sel = np.concatenate([flat[starts[L]:ends[L]] for L in np.flatnonzero(tm.ravel())])
    S = np.unique(sel)
    inv = np.zeros(N, np.int32)
    inv[S] = np.arange(S.size, dtype=np.int32)
    struct = (meta["tiles_per_gauss"].cpu().numpy()[:, S], meta["isect_ids"].cpu().numpy(), inv[flat], offs[None])
    return S, struct


def test_full_size_batch_gradients_vs_summed_fp64_oracle(device):
    """The batched backward against the FP64 oracle: the loss is each view's
    L1 over 64 random tiles (tile seed 1 + view), summed over the 8 views as
    the reference's per-view loop accumulates it (train.py:108-128); every
    leaf gradient of one render_batch backward must match the sum over views
    of the oracle's FP64 beta_step (blending those tiles in the view's own
    tile lists) to a relative L2 of 1e-3.  The oracle runs each view on the
    primitives its 64 tiles reach; the others have zero gradient in both."""
    from oracle import ubs_oracle as O

    sizes = [SIZES[c % len(SIZES)] for c in range(VIEWS)]
    sc = scenes.make_scene(N, D, seed=0)
    p = beta.BetaParams.from_scene(sc, device)
    vms, _, _ = scenes.make_cameras(VIEWS, sizes[0][0], sizes[0][1], seed=0)
    Ks = np.stack([scenes.intrinsics(w, h) for w, h in sizes]).astype(np.float32)
    vms_t, Ks_t = torch.tensor(vms, device=device), torch.tensor(Ks, device=device)
    g = np.random.Generator(np.random.PCG64(5))
    tgts = [g.uniform(0, 1, (h, w, 3)) for w, h in sizes]
    tms, pms = [], []
    for c, (w, h) in enumerate(sizes):
        tw, th = -(-w // 16), -(-h // 16)
        m = np.zeros(th * tw, bool)
        m[np.random.Generator(np.random.PCG64(1 + c)).choice(th * tw, 64, replace=False)] = True
        tms.append(m.reshape(th, tw))
        pms.append(np.kron(tms[-1], np.ones((16, 16), bool))[:h, :w])
    print(f"C4 batch vs FP64: N={N} views={sizes}", flush=True)
    imgs, _ = beta.render_batch(p, vms_t, Ks_t, sizes)
    loss = 0.0
    for im, t, pm in zip(imgs, tgts, pms):
        pm_t = torch.tensor(pm, device=device)[..., None]
        loss = loss + ((im - torch.tensor(t, dtype=torch.float32, device=device)).abs() * pm_t).sum() / (
            pm.sum() * 3)
    loss.backward()
    leaves = {k: getattr(sc, k) for k in ("xyz", "mean_rest", "scale_raw", "l_triangle", "opacity_raw", "beta_raw",
                                         "rgb", "quats")}
    want = {k: np.zeros(v.shape, np.float64) for k, v in leaves.items()}
    reached = 0
    for c, (w, h) in enumerate(sizes):
        with torch.no_grad():
            _, _, meta = beta.render_views(p, vms_t[c:c + 1], Ks_t[c:c + 1], w, h)
        S, struct = _view_subset(meta, tms[c], N)
        reached += S.size
        _, gc, _ = O.beta_step({k: np.ascontiguousarray(v[S]) for k, v in leaves.items()}, vms[c], Ks[c], w, h,
                               tgts[c], dtype=np.float64, structure=struct, tile_mask=tms[c])
        for k in want:
            want[k][S] += gc[k]
        print(f"view {c} {w}x{h}: {S.size} primitives reach its 64 tiles", flush=True)
    assert reached > 8 * 10_000
    for k, t in p.leaves().items():
        if t is None:
            continue
        r = rel_l2(t.grad.cpu().numpy(), want[k])
        print(f"{k}: rel-L2 {r:.2e}", flush=True)
        assert r < 1e-3, (k, r)


def test_full_size_batch_checkpointed_transmittance(device):
    """The batched step with the per-batch transmittance checkpoints (the
    forward stores T after every batch of 256 entries, the backward restarts
    each batch from it) against the same step recomputing T by division:
    images bit-identical, every leaf gradient within a relative L2 of 1e-4
    (the two recover T before a primitive through different roundings)."""
    from gsplat.cuda import _wrapper as W

    sizes = [SIZES[c % len(SIZES)] for c in range(VIEWS)]
    sc = scenes.make_scene(N, D, seed=0)
    vms, _, _ = scenes.make_cameras(VIEWS, sizes[0][0], sizes[0][1], seed=0)
    Ks = np.stack([scenes.intrinsics(w, h) for w, h in sizes]).astype(np.float32)
    vms_t, Ks_t = torch.tensor(vms, device=device), torch.tensor(Ks, device=device)
    g = np.random.Generator(np.random.PCG64(5))
    tgts = [torch.tensor(g.uniform(0, 1, (h, w, 3)), dtype=torch.float32, device=device) for w, h in sizes]
    out = {}
    for mode in (False, True):
        p = beta.BetaParams.from_scene(sc, device)
        W.set_backward_transmittance_checkpoint(mode)
        try:
            imgs, _ = beta.render_batch(p, vms_t, Ks_t, sizes)
            sum((im - t).abs().mean() for im, t in zip(imgs, tgts)).backward()
        finally:
            W.set_backward_transmittance_checkpoint(False)
        out[mode] = ([im.detach() for im in imgs], {k: v.grad.detach().clone() for k, v in p.leaves().items()})
        del p
    print(f"C4 checkpointed vs recomputed transmittance: N={N} views={sizes}", flush=True)
    for a, b in zip(out[False][0], out[True][0]):
        assert torch.equal(a, b)
    for k, a in out[False][1].items():
        b = out[True][1][k]
        rel = ((a.double() - b.double()).norm() / a.double().norm()).item()
        print(f"{k}: rel-L2 {rel:.2e}", flush=True)
        assert rel <= 1e-4, (k, rel)
