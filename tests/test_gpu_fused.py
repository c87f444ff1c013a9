"""GPU parity of the fused Beta projection (ubs_beta_project_fwd/bwd,
gsplat.beta.render_views) against the drop-in chain it fuses.

The drop-in chain (activations -> rotation -> Sigma = L L^T -> conditioning
-> covars33 projection) is itself pinned bit for bit / to tolerance against
the oracle by test_gpu_kernels.py and test_gpu_e2e.py.  The fused kernels call
the same device functions in the same order, so with one view:
* forward outputs (radii; means2d, depths, conics of visible entries;
  conditioned opacities, betas) are bit-identical;
* every leaf gradient is bit-identical (compared by value: torch.equal, so
  +0 and -0 are the same zero).
With several views the per-view forward is still bit-identical to running the
chain per view, and the leaf gradients equal the sum of the per-view chains'
within rtol 1e-4 (the fused kernel sums the views' gradients before the
covariance / rotation / activation VJPs, the chain after them).
"""
import numpy as np
import pytest
import torch

from gsplat import beta, scenes
from gsplat.cuda import _wrapper as W

pytestmark = pytest.mark.gpu

LEAVES = ("xyz", "mean_rest", "scale_raw", "l_triangle", "opacity_raw", "beta_raw", "quats")


def _params(device, N, D, quats, seed):
    sc = scenes.make_scene(N, D, seed=seed)
    if not quats:
        # give the first-order rotation non-zero skew parameters
        g = np.random.Generator(np.random.PCG64(seed + 7))
        sc.l_triangle[:, :3] = g.normal(0.0, 0.3, (N, 3)).astype(np.float32)
    return beta.BetaParams.from_scene(sc, device, requires_grad=True, use_quats=quats)


def _leaves(p):
    return [getattr(p, k) for k in LEAVES if getattr(p, k) is not None]


def _chain(p, vm, K, Wd, Ht, t):
    """The drop-in path of one view: conditioned() + covars33 projection."""
    m, V, o, b0, _ = beta.conditioned(p, vm, t)
    radii, means2d, depths, conics, _ = W.fully_fused_projection_covars33(m, V, vm[None], K[None], Wd, Ht)
    return radii, means2d, depths, conics, o[None], b0[None]


def _fused(p, vms, Ks, Wd, Ht, ts, image_sizes=None):
    return W.beta_project(p.xyz, p.mean_rest, p.scale_raw, p.l_triangle, p.opacity_raw, p.beta_raw, p.quats,
                          vms, Ks, Wd, Ht, cam_pos=W.camera_centers(vms),
                          timestamps=None if ts is None else torch.as_tensor(ts, dtype=torch.float32,
                                                                             device=vms.device),
                          image_sizes=image_sizes)


def _upstream(radii, seed):
    C, N = radii.shape
    g = torch.Generator(device="cpu").manual_seed(seed)
    vis = (radii > 0).to(torch.float32).cpu()
    out = [torch.randn((C, N, 2), generator=g), torch.randn((C, N), generator=g),
           torch.randn((C, N, 3), generator=g) * 1e-2, torch.randn((C, N), generator=g) * vis,
           torch.randn((C, N), generator=g) * vis]
    return [x.to(radii.device) for x in out]


def _loss(outs, ups):
    return sum((o * u).sum() for o, u in zip(outs[1:], ups))


def _cams(device, C, Wd, Ht, seed):
    vms, Ks, ts = scenes.make_cameras(C, Wd, Ht, seed=seed)
    return torch.tensor(vms, device=device), torch.tensor(Ks, device=device), ts


@pytest.mark.parametrize("D,quats", [(6, True), (6, False), (7, True), (7, False)])
def test_fused_projection_bit_identical_to_chain(device, D, quats):
    N, Wd, Ht = 6000, 256, 192
    p = _params(device, N, D, quats, seed=D + 10 * quats)
    vms, Ks, ts = _cams(device, 1, Wd, Ht, seed=3)
    t = float(ts[0]) if D == 7 else None
    ref = _chain(p, vms[0], Ks[0], Wd, Ht, t)
    got = _fused(p, vms, Ks, Wd, Ht, [t] if t is not None else None)
    assert torch.equal(got[0], ref[0]), "radii"
    vis = ref[0] > 0
    assert int(vis.sum()) > N // 4, "too few visible primitives for a meaningful test"
    for k, name in ((1, "means2d"), (2, "depths"), (3, "conics")):
        assert torch.equal(got[k][vis], ref[k][vis]), name
    assert torch.equal(got[4], ref[4]), "opacities"
    assert torch.equal(got[5], ref[5]), "betas"
    ups = _upstream(ref[0], seed=5)
    g_ref = torch.autograd.grad(_loss(ref, ups), _leaves(p))
    g_got = torch.autograd.grad(_loss(got, ups), _leaves(p))
    for a, b, name in zip(g_got, g_ref, [k for k in LEAVES if getattr(p, k) is not None]):
        assert torch.equal(a, b), (name, (a - b).abs().max().item())


@pytest.mark.parametrize("D,C", [(6, 3), (7, 4)])
def test_fused_projection_multiview(device, D, C):
    N, Wd, Ht = 5000, 224, 160
    p = _params(device, N, D, True, seed=21 + D)
    vms, Ks, ts = _cams(device, C, Wd, Ht, seed=4)
    tl = [float(x) for x in ts] if D == 7 else None
    got = _fused(p, vms, Ks, Wd, Ht, tl)
    ups = _upstream(got[0], seed=9)
    g_got = torch.autograd.grad(_loss(got, ups), _leaves(p))
    g_ref = [torch.zeros_like(x) for x in _leaves(p)]
    for c in range(C):
        ref = _chain(p, vms[c], Ks[c], Wd, Ht, tl[c] if tl else None)
        assert torch.equal(got[0][c], ref[0][0]), ("radii", c)
        vis = ref[0][0] > 0
        for k in (1, 2, 3):
            assert torch.equal(got[k][c][vis], ref[k][0][vis]), (k, c)
        assert torch.equal(got[4][c], ref[4][0]) and torch.equal(got[5][c], ref[5][0]), c
        gc = torch.autograd.grad(_loss(ref, [u[c:c + 1] for u in ups]), _leaves(p))
        g_ref = [a + b for a, b in zip(g_ref, gc)]
    for a, b, name in zip(g_got, g_ref, [k for k in LEAVES if getattr(p, k) is not None]):
        scale = b.abs().max().item()
        torch.testing.assert_close(a, b, rtol=1e-4, atol=1e-5 * scale, msg=name)


def test_fused_projection_per_view_sizes(device):
    N, D = 4000, 6
    p = _params(device, N, D, True, seed=31)
    sizes = [(320, 180), (160, 120), (256, 256)]
    vms, Ks, _ = _cams(device, 3, 320, 180, seed=6)
    Ks = Ks.clone()
    for c, (w, h) in enumerate(sizes):  # principal point at each view's centre
        Ks[c, 0, 2], Ks[c, 1, 2] = w / 2.0, h / 2.0
    wh = torch.tensor(sizes, dtype=torch.int32, device=device)
    got = _fused(p, vms, Ks, 0, 0, None, image_sizes=wh)
    for c, (w, h) in enumerate(sizes):
        one = _fused(p, vms[c:c + 1], Ks[c:c + 1], w, h, None)
        assert torch.equal(got[0][c], one[0][0]), c
        vis = one[0][0] > 0
        for k in (1, 2, 3, 4, 5):
            a, b = got[k][c], one[k][0]
            if k <= 3:
                a, b = a[vis], b[vis]
            assert torch.equal(a, b), (k, c)


def test_fused_projection_empty(device):
    p = _params(device, 0, 6, True, seed=1)
    vms, Ks, _ = _cams(device, 2, 64, 64, seed=1)
    out = _fused(p, vms, Ks, 64, 64, None)
    assert [tuple(o.shape) for o in out] == [(2, 0), (2, 0, 2), (2, 0), (2, 0, 3), (2, 0), (2, 0)]


@pytest.mark.parametrize("D", [6, 7])
def test_render_views_matches_render(device, D):
    """Images bit-identical; leaf gradients to rtol 1e-4, as the blend's
    float atomics add a primitive's per-tile gradients in a run-dependent
    order (two runs of render() itself differ in the last bits)."""
    N, Wd, Ht = 20_000, 256, 256
    p = _params(device, N, D, True, seed=41 + D)
    vms, Ks, ts = _cams(device, 1, Wd, Ht, seed=7)
    t = float(ts[0]) if D == 7 else None
    img_r, a_r, _ = beta.render(p, vms[0], Ks[0], Wd, Ht, timestamp=t)
    img_f, a_f, meta = beta.render_views(
        p, vms, Ks, Wd, Ht, timestamps=None if t is None else torch.tensor([t], device=device))
    assert torch.equal(img_f, img_r) and torch.equal(a_f, a_r)
    assert meta["isect_ids"].numel() > 0
    tgt = torch.rand(img_r.shape, generator=torch.Generator().manual_seed(2)).to(device)
    names = [k for k in LEAVES if getattr(p, k) is not None] + ["rgb"]
    g_f = torch.autograd.grad((img_f - tgt).abs().mean(), _leaves(p) + [p.rgb])
    g_r = torch.autograd.grad((img_r - tgt).abs().mean(), _leaves(p) + [p.rgb])
    for a, b, name in zip(g_f, g_r, names):
        scale = b.abs().max().item()
        torch.testing.assert_close(a, b, rtol=1e-4, atol=1e-5 * scale, msg=name)


def test_render_views_multiview_matches_per_view(device):
    N, Wd, Ht, C = 20_000, 192, 160, 3
    p = _params(device, N, 6, True, seed=51)
    vms, Ks, _ = _cams(device, C, Wd, Ht, seed=8)
    img_f, a_f, _ = beta.render_views(p, vms, Ks, Wd, Ht)
    tgt = torch.rand(img_f.shape, generator=torch.Generator().manual_seed(3)).to(device)
    g_f = torch.autograd.grad((img_f - tgt).abs().sum(), _leaves(p) + [p.rgb])
    g_r = None
    for c in range(C):
        img_c, a_c, _ = beta.render(p, vms[c], Ks[c], Wd, Ht)
        assert torch.equal(img_f[c], img_c[0]) and torch.equal(a_f[c], a_c[0]), c
        gc = torch.autograd.grad((img_c - tgt[c:c + 1]).abs().sum(), _leaves(p) + [p.rgb])
        g_r = list(gc) if g_r is None else [x + y for x, y in zip(g_r, gc)]
    for a, b, name in zip(g_f, g_r, [k for k in LEAVES if getattr(p, k) is not None] + ["rgb"]):
        scale = b.abs().max().item()
        # the blend's float atomics add the views' colour gradients in another order
        torch.testing.assert_close(a, b, rtol=1e-4, atol=1e-5 * scale, msg=name)


def test_render_batch_mixed_sizes_matches_single_views(device):
    """render_batch over views of three sizes: every image bit-identical to
    rendering that view alone with render_views; leaf gradients of the
    summed loss equal the per-view sums to rtol 1e-4."""
    N = 20_000
    p = _params(device, N, 6, True, seed=61)
    sizes = [(160, 96), (224, 128), (256, 160), (160, 96), (224, 128)]
    vms, Ks, _ = _cams(device, len(sizes), 256, 160, seed=9)
    Ks = Ks.clone()
    for c, (w, h) in enumerate(sizes):
        Ks[c, 0, 0] = Ks[c, 1, 1] = w / (2 * 0.57735026919)
        Ks[c, 0, 2], Ks[c, 1, 2] = w / 2.0, h / 2.0
    imgs, alphas = beta.render_batch(p, vms, Ks, sizes)
    gen = torch.Generator().manual_seed(4)
    tgts = [torch.rand((h, w, 3), generator=gen).to(device) for w, h in sizes]
    loss = sum((im - t).abs().sum() for im, t in zip(imgs, tgts))
    g_b = torch.autograd.grad(loss, _leaves(p) + [p.rgb])
    g_r = None
    for c, (w, h) in enumerate(sizes):
        one, a1, _ = beta.render_views(p, vms[c:c + 1], Ks[c:c + 1], w, h)
        assert torch.equal(imgs[c], one[0]) and torch.equal(alphas[c], a1[0]), c
        gc = torch.autograd.grad((one[0] - tgts[c]).abs().sum(), _leaves(p) + [p.rgb])
        g_r = list(gc) if g_r is None else [x + y for x, y in zip(g_r, gc)]
    for a, b, name in zip(g_b, g_r, [k for k in LEAVES if getattr(p, k) is not None] + ["rgb"]):
        scale = b.abs().max().item()
        torch.testing.assert_close(a, b, rtol=1e-4, atol=1e-5 * scale, msg=name)


def test_ply_scene_renders_like_the_in_memory_scene(device, tmp_path):
    """A scene written in the reference's PLY layout and read back into HBM
    renders bit-identically to the same leaves built in memory."""
    p = _params(device, 5000, 7, False, seed=71)
    path = str(tmp_path / "scene.ply")
    p.save_ply(path)
    q = beta.BetaParams.from_ply(path, device)
    vms, Ks, ts = _cams(device, 2, 192, 128, seed=10)
    tt = torch.tensor([float(x) for x in ts], device=device)
    a, _, _ = beta.render_views(p, vms, Ks, 192, 128, timestamps=tt)
    b, _, _ = beta.render_views(q, vms, Ks, 192, 128, timestamps=tt)
    assert torch.equal(a, b)
