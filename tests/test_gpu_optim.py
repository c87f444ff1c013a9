"""GPU parity of the step after the backward (ubs_adam_step, ubs_xyz_noise,
ubs_mcmc_relocate through gsplat.optim) against torch.optim.Adam on the GPU
(what the reference runs) and the fp64 oracle (pinned on the CPU by
tests/test_optim_host.py).

Tolerances: Adam parameters within 4 fp32 ulps of |p| plus 1e-6 lr per step
of each other (torch's kernels may contract m, v updates into FMAs; ours
keeps the unfused order), moments to 1e-6 relative plus 4 ulps of the
largest moment (sums of gradients of mixed scales cancel); noise 1e-5 of its
largest component (fp32 covariance products); relocated rows bit-exact,
their split opacity logit within 2e-5 (fp32 pow / log against fp64).
"""
import numpy as np
import pytest
import torch

from gsplat import optim as G
from gsplat import scenes
from gsplat.beta import BetaParams
from oracle import ubs_oracle as O

pytestmark = pytest.mark.gpu


def _adam_pair(device, shapes, lrs, misalign, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    ps = []
    for s, mis in zip(shapes, misalign):
        n = int(np.prod(s))
        if mis:  # a contiguous view 4 bytes into its buffer: the scalar path
            buf = torch.randn(n + 1, generator=g).to(device)
            ps.append(buf[1:].view(s))
        else:
            ps.append(torch.randn(s, generator=g).to(device))
    return ps, [p.detach().clone() for p in ps]


@pytest.mark.parametrize("seed", [0, 1])
def test_adam_matches_torch_adam(device, seed):
    shapes = [(1000, 3), (1000, 0), (1000, 6), (1000, 15), (1000, 1), (7,), (2049, 4), (100_003,)]
    lrs = [1.6e-4, 1e-3, 5e-3, 1e-3, 0.05, 1e-3, 1e-3, 0.01]
    misalign = [False, False, False, True, False, True, False, seed == 1]
    ours, ref = _adam_pair(device, shapes, lrs, misalign, seed)
    ref = [torch.nn.Parameter(t) for t in ref]
    topt = torch.optim.Adam([{"params": [t], "lr": lr} for t, lr in zip(ref, lrs)], lr=0.0, eps=1e-15)
    names = [f"t{i}" for i in range(len(shapes))]

    class P:  # a parameter holder with the fields BetaAdam reads
        pass
    holder = P()
    for n, t in zip(names, ours):
        t.requires_grad_(True)
        setattr(holder, n, t)
    opt = G.BetaAdam(holder, dict(zip(names, lrs)))
    gen = torch.Generator(device="cpu").manual_seed(100 + seed)
    steps = 5
    for step in range(steps):
        for k, (a, b) in enumerate(zip(ours, ref)):
            if step == 2 and k == 2:  # a leaf without a gradient is skipped, its step count kept
                a.grad, b.grad = None, None
                continue
            gr = (torch.randn(a.shape, generator=gen) * 10.0 ** float(torch.randint(-7, 1, (1,), generator=gen)))
            a.grad = gr.to(device)
            b.grad = gr.to(device).clone()
        opt.step()
        topt.step()
    torch.cuda.synchronize()
    assert opt.steps["t2"] == steps - 1 and opt.steps["t0"] == steps
    for k, (a, b) in enumerate(zip(ours, ref)):
        st = topt.state[b] if b.numel() else None
        a_, b_ = a.detach().cpu().numpy(), b.detach().cpu().numpy()
        tol = 4 * np.finfo(np.float32).eps * np.abs(b_) + 1e-6 * lrs[k] * steps
        assert np.all(np.abs(a_ - b_) <= tol), (k, np.abs(a_ - b_).max())
        if st is not None:  # moments: cancellation across steps, so ulps of the largest term
            m_ref, v_ref = st["exp_avg"].cpu().numpy(), st["exp_avg_sq"].cpu().numpy()
            np.testing.assert_allclose(opt.exp_avg[names[k]].cpu().numpy(), m_ref, rtol=1e-6,
                                       atol=4 * np.finfo(np.float32).eps * np.abs(m_ref).max())
            np.testing.assert_allclose(opt.exp_avg_sq[names[k]].cpu().numpy(), v_ref, rtol=1e-6,
                                       atol=4 * np.finfo(np.float32).eps * np.abs(v_ref).max())


def test_adam_matches_fp64_oracle(device):
    rng = np.random.Generator(np.random.PCG64(9))
    p0 = rng.normal(0, 1, (4096, 5)).astype(np.float32)
    t = torch.tensor(p0, device=device, requires_grad=True)

    class P:
        pass
    h = P()
    h.x = t
    opt = G.BetaAdam(h, {"x": 0.005})
    p, m, v = p0.astype(np.float64), np.zeros(p0.shape), np.zeros(p0.shape)
    for step in range(1, 4):
        g = (rng.normal(0, 1, p0.shape) * 1e-3).astype(np.float32)
        t.grad = torch.tensor(g, device=device)
        opt.step()
        p, m, v = O.adam_step(p, g, m, v, 0.005, step)
    got = t.detach().cpu().numpy()
    assert np.abs(got - p).max() <= 1e-6 * np.abs(p).max() + 1e-7 * 3, np.abs(got - p).max()


def test_beta_adam_for_model_groups(device):
    """BetaAdam.for_model on a scene: the seven reference groups with their
    learning rates, and the xyz rate following the schedule."""
    sc = scenes.make_scene(200, 6, seed=4)
    params = BetaParams.from_scene(sc, device, use_quats=False)
    opt = G.BetaAdam.for_model(params, spatial_lr_scale=2.0)
    assert set(opt.exp_avg) == {f for _, f in G.LEAF_GROUPS}
    assert opt.lrs["opacity_raw"] == 0.05 and opt.lrs["scale_raw"] == 0.005
    assert opt.update_learning_rate(0) == pytest.approx(0.00032, rel=1e-12)
    before = {f: getattr(params, f).detach().clone() for f in opt.exp_avg}
    for f in opt.exp_avg:
        getattr(params, f).grad = torch.ones_like(getattr(params, f))
    opt.step()
    for f in opt.exp_avg:  # step 1: p' = p - lr * sign(g) (|g| >> eps)
        want = before[f] - opt.lrs[f]
        torch.testing.assert_close(getattr(params, f).detach(), want, rtol=0, atol=4e-7 * (1 + before[f].abs().max()))
    opt.zero_grad()
    assert all(getattr(params, f).grad is None for f in opt.exp_avg)


@pytest.mark.parametrize("D", [6, 7])
def test_xyz_noise_matches_oracle(device, D):
    sc = scenes.make_scene(5000, D, seed=11)
    rng = np.random.Generator(np.random.PCG64(D))
    sc.l_triangle[:, :3] = rng.normal(0, 0.3, (5000, 3))  # a non-zero skew: R = I + A
    sc.opacity_raw -= 3.0  # enough translucent primitives for (1 - o)^100 to matter
    params = BetaParams.from_scene(sc, device, requires_grad=False, use_quats=False)
    xyz0 = params.xyz.detach().cpu().numpy()
    z = torch.randn((5000, 3), generator=torch.Generator().manual_seed(2)).to(device)
    xyz_lr = 1.6e-4
    G.add_xyz_noise(params, xyz_lr, noise_lr=5e3, normal=z)
    torch.cuda.synchronize()
    want = O.xyz_noise(xyz0.astype(np.float64), sc.l_triangle, sc.scale_raw, sc.opacity_raw, z.cpu().numpy(),
                       5e3, xyz_lr)
    got = params.xyz.cpu().numpy()
    d = np.abs(want - xyz0).max()
    assert d > 1e-5  # the noise moved something
    assert np.abs(got - want).max() <= 1e-5 * d + 2 * np.finfo(np.float32).eps * np.abs(want).max()
    # the reference's own expression (train.py:159-168) on the GPU, fp32
    o = torch.sigmoid(torch.tensor(sc.opacity_raw, device=device))
    S = torch.tensor(O.xyz_covariance(sc.l_triangle, sc.scale_raw), dtype=torch.float32, device=device)
    ref = torch.tensor(xyz0, device=device) + torch.bmm(S, (z * torch.pow(1 - o, 100) * 5e3 * xyz_lr)[..., None])[..., 0]
    assert float((params.xyz - ref).abs().max()) <= 1e-5 * d + 2 * np.finfo(np.float32).eps * np.abs(want).max()


def _scene_params(device, N, D, seed):
    sc = scenes.make_scene(N, D, seed=seed)
    return sc, BetaParams.from_scene(sc, device, use_quats=False)


def test_mcmc_relocate_rows_match_oracle(device):
    sc, params = _scene_params(device, 3000, 7, 12)
    opt = G.BetaAdam.for_model(params)
    gen = torch.Generator().manual_seed(3)
    for f in opt.exp_avg:  # nonzero moments so that the resets are visible
        opt.exp_avg[f].copy_(torch.randn(opt.exp_avg[f].shape, generator=gen))
        opt.exp_avg_sq[f].copy_(torch.rand(opt.exp_avg_sq[f].shape, generator=gen))
    leaves0 = {f: getattr(params, f).detach().cpu().numpy().copy() for _, f in G.LEAF_GROUPS}
    m0 = {f: opt.exp_avg[f].cpu().numpy().copy() for f in opt.exp_avg}
    v0 = {f: opt.exp_avg_sq[f].cpu().numpy().copy() for f in opt.exp_avg}
    perm = torch.randperm(3000, generator=gen)
    dst = perm[:300]
    src = perm[300:][torch.randint(0, 2700, (300,), generator=gen)]
    ratio = torch.bincount(src)[src]
    G._relocate_rows(opt, params, src.to(device), dst.to(device), ratio.to(device))
    torch.cuda.synchronize()
    L, M, V = O.mcmc_relocate(leaves0, m0, v0, src.numpy(), dst.numpy(), ratio.numpy())
    for _, f in G.LEAF_GROUPS:
        got = getattr(params, f).detach().cpu().numpy()
        if f == "opacity_raw":
            np.testing.assert_allclose(got, L[f], rtol=0, atol=2e-5)
            np.testing.assert_array_equal(got[dst.numpy()], got[src.numpy()])
        else:
            np.testing.assert_array_equal(got, L[f].astype(np.float32))
        np.testing.assert_array_equal(opt.exp_avg[f].cpu().numpy(), M[f].astype(np.float32))
        np.testing.assert_array_equal(opt.exp_avg_sq[f].cpu().numpy(), V[f].astype(np.float32))


def test_relocate_and_add_new_gs(device):
    """relocate_gs moves every dead primitive onto a drawn alive one;
    add_new_gs grows by 2% with zero moments in the new rows; a render and a
    fused Adam step run on the grown scene."""
    from gsplat.beta import render_views

    sc, params = _scene_params(device, 4000, 6, 13)
    with torch.no_grad():
        params.opacity_raw[:500] = -8.0  # opacity 3e-4 <= 0.005: dead
    opt = G.BetaAdam.for_model(params)
    dead = (torch.sigmoid(params.opacity_raw) <= 0.005).squeeze(-1)
    gen = torch.Generator(device=device).manual_seed(5)
    moved = G.relocate_gs(params, opt, dead, gen)
    assert moved == 500
    assert int((torch.sigmoid(params.opacity_raw) <= 0.005).sum()) == 0
    lt = params.l_triangle.detach()
    # each relocated row equals some alive row in l_triangle
    assert bool((lt[:500, None, :] == lt[None, 500:, :]).all(-1).any(1).all())
    added = G.add_new_gs(params, opt, cap_max=10_000, generator=gen)
    assert added == 80 and params.xyz.shape[0] == 4080
    for f in opt.exp_avg:
        assert opt.exp_avg[f].shape[0] == 4080 and getattr(params, f).shape[0] == 4080
        assert not bool(opt.exp_avg[f][4000:].any()) and not bool(opt.exp_avg_sq[f][4000:].any())
    assert G.add_new_gs(params, opt, cap_max=4080, generator=gen) == 0
    vm = torch.tensor(scenes.look_at_viewmat(np.array([0.0, 0.3, 4.0])), dtype=torch.float32, device=device)[None]
    K = torch.tensor(scenes.intrinsics(128, 96), dtype=torch.float32, device=device)[None]
    img, _, _ = render_views(params, vm, K, 128, 96)
    img.sum().backward()
    opt.step()
    assert all(torch.isfinite(getattr(params, f)).all() for f in opt.exp_avg)
