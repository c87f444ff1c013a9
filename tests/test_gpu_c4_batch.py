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
