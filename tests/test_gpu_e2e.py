"""End-to-end GPU parity of the drop-in path (BetaModel.render mirror +
rasterization()) against the oracle.

* Structure (covariance, conditioned mean/covariance, projection, intersection
  lists, sort order, tile ranges) is bit-exact given the same fp32 inputs.
* Image: max-abs <= 1e-4 against the FP64 oracle blending the same tile lists,
  on pixels away from the transmittance cutoff branch point.
* Leaf-parameter gradients of the L1 loss: relative L2 <= 1e-3 (fp32 atomics
  and transmittance recovery against FP64).
"""
import numpy as np
import pytest
import torch

from gsplat import beta, rasterization, scenes
from gsplat.cuda import _wrapper as W
from oracle import ubs_oracle as O

from util import assert_same, check_branch_pixels, rel_l2

pytestmark = pytest.mark.gpu


def host(t):
    return t.detach().cpu().numpy()


def _leaves(sc):
    return dict(xyz=sc.xyz, mean_rest=sc.mean_rest, scale_raw=sc.scale_raw, l_triangle=sc.l_triangle,
                opacity_raw=sc.opacity_raw, beta_raw=sc.beta_raw, rgb=sc.rgb, quats=sc.quats)


@pytest.mark.parametrize("D,N,Wd,Ht", [(6, 10_000, 256, 256), (7, 8_000, 200, 152)])
def test_structure_bit_exact_chain(device, D, N, Wd, Ht):
    sc = scenes.make_scene(N, D, seed=1)
    vms, Ks, ts = scenes.make_cameras(1, Wd, Ht, seed=1)
    p = beta.BetaParams.from_scene(sc, device, requires_grad=False)
    vm = torch.tensor(vms[0], device=device)
    K = torch.tensor(Ks[0], device=device)
    # GPU activations and rotation, then every gsplat op on GPU and on the oracle
    scale = torch.nn.functional.softplus(p.scale_raw)
    opac = torch.sigmoid(p.opacity_raw)
    bet = 4.0 * torch.exp(p.beta_raw)
    rot = beta.quat_to_rotmat(p.quats)
    ri, rj = beta.rest_indices(D, device)
    cov = W.rot_scale_l_triangle_to_covar(rot, scale, p.l_triangle, ri, rj)
    assert_same(host(cov), O.covar_fwd(host(rot), host(scale), host(p.l_triangle), host(ri), host(rj)), "covar")
    mean = torch.cat([p.xyz, p.mean_rest], -1)
    cc = beta.camera_center(vm)
    vd = p.xyz - cc[None]
    vd = vd / vd.norm(dim=-1, keepdim=True)
    q = vd if D == 6 else torch.cat([vd, torch.full_like(vd[:, :1], float(ts[0]))], -1)
    m, V, o = W.cond_mean_convariance_opacity(mean, cov, opac, bet[:, 1:].contiguous(), q)
    om, oV, oo = O.cond_fwd(host(mean), host(cov), host(opac), host(bet[:, 1:]), host(q))
    assert_same(host(m), om, "m_cond")
    assert_same(host(V), oV, "V_cond")
    img, alpha, meta = rasterization(m, p.l_triangle, scale, o[:, 0], bet[:, 0].contiguous(), p.rgb, vm[None], K[None],
                                     Wd, Ht, covars=V)
    ref_img, ref_a, rm = O.rasterization(host(m), host(V), host(o[:, 0]), host(bet[:, 0]), host(p.rgb),
                                         vms[:1], Ks[:1], Wd, Ht)
    for k in ("radii", "tiles_per_gauss", "isect_ids", "flatten_ids", "isect_offsets"):
        assert_same(host(meta[k]), rm[k], k)
    vis = rm["radii"] > 0
    for k in ("means2d", "depths", "conics"):
        assert_same(host(meta[k])[vis], rm[k][vis], k)
    assert rm["isect_ids"].size > 1000
    np.testing.assert_allclose(host(img), ref_img, atol=5e-5)


@pytest.mark.parametrize("D", [6, 7])
def test_step_image_and_gradients_vs_fp64(device, D):
    N, Wd, Ht = (6_000, 160, 128) if D == 6 else (5_000, 144, 112)
    sc = scenes.make_scene(N, D, seed=2)
    vms, Ks, ts = scenes.make_cameras(1, Wd, Ht, seed=2)
    tgt = np.random.Generator(np.random.PCG64(7)).uniform(0, 1, (Ht, Wd, 3))
    p = beta.BetaParams.from_scene(sc, device)
    img, alpha, meta = beta.render(p, torch.tensor(vms[0], device=device), torch.tensor(Ks[0], device=device),
                                   Wd, Ht, timestamp=float(ts[0]))
    loss = (img[0] - torch.tensor(tgt, dtype=torch.float32, device=device)).abs().mean()
    loss.backward()
    # the FP64 oracle blends in the GPU's tile lists (their bit-exactness is
    # test_structure_bit_exact_chain's job)
    struct = tuple(host(meta[k]) for k in ("tiles_per_gauss", "isect_ids", "flatten_ids", "isect_offsets"))
    ref_img, ref_g, rm = O.beta_step(_leaves(sc), vms[0], Ks[0], Wd, Ht, tgt, timestamp=float(ts[0]),
                                     dtype=np.float64, structure=struct)
    _, _, _, mT, mS = O.raster_fwd(rm["means2d"], rm["conics"], rm["colors"], rm["opacities"], rm["betas"], Wd, Ht,
                                   16, rm["isect_offsets"], rm["flatten_ids"], margins=True)
    safe = (mT[0] > 1e-3) & (mS[0] > 1e-6)
    assert safe.mean() > 0.99
    err = np.abs(host(img[0]) - ref_img)[safe].max()
    assert err <= 1e-4, err
    g32 = {k: host(meta[k]) for k in ("means2d", "conics", "opacities", "betas", "isect_offsets", "flatten_ids")}
    oc, _, olast = O.raster_fwd(g32["means2d"], g32["conics"], host(p.rgb)[None], g32["opacities"], g32["betas"],
                                Wd, Ht, 16, g32["isect_offsets"], g32["flatten_ids"])
    with torch.no_grad():
        _, _, last, _ = W.raster_fwd_raw(meta["means2d"], meta["conics"], p.rgb[None].contiguous(),
                                         meta["opacities"], meta["betas"], None, None, Wd, Ht, 16,
                                         meta["isect_offsets"], meta["flatten_ids"])
    check_branch_pixels(host(img[0]), host(last[0]), oc[0], olast[0], np.ones(safe.shape, bool), safe,
                        what=f"{D}-D step near branch points")
    for k, t in p.leaves().items():
        r = rel_l2(host(t.grad), ref_g[k])
        assert r < 1e-3, (k, r)


def test_render_modes_and_options(device):
    N, Wd, Ht = 3000, 128, 96
    sc = scenes.make_scene(N, 6, seed=4)
    vms, Ks, ts = scenes.make_cameras(2, Wd, Ht, seed=4)
    p = beta.BetaParams.from_scene(sc, device, requires_grad=False)
    m, V, o, b0, scale = beta.conditioned(p, torch.tensor(vms[0], device=device))
    vmt, Kt = torch.tensor(vms, device=device), torch.tensor(Ks, device=device)
    bg = torch.rand(2, 3, device=device)
    kw = dict(means=m, l_triagnles=p.l_triangle, scales=scale, opacities=o, betas=b0, colors=p.rgb, viewmats=vmt,
              Ks=Kt, width=Wd, height=Ht, covars=V)
    rgb, a, meta = rasterization(**kw, backgrounds=bg)
    assert rgb.shape == (2, Ht, Wd, 3) and a.shape == (2, Ht, Wd, 1)
    dep, a2, _ = rasterization(**kw, render_mode="Depth")
    rgbd, a3, _ = rasterization(**kw, render_mode="RGB+D", backgrounds=bg)
    assert_same(host(rgbd[..., :3]), host(rgb), "RGB part of RGB+D")
    np.testing.assert_allclose(host(rgbd[..., 3:]), host(dep), rtol=1e-6, atol=1e-6)
    ed, _, _ = rasterization(**kw, render_mode="RGB+ED")
    np.testing.assert_allclose(host(ed[..., 3:]), host(dep / a2.clamp(min=1e-10)), rtol=1e-5, atol=1e-6)
    # channel chunking (colors wider than channel_chunk) and odd channel counts
    feats = torch.rand(N, 37, device=device)
    f1, _, _ = rasterization(**{**kw, "colors": feats}, channel_chunk=32)
    f2, _, _ = rasterization(**{**kw, "colors": feats}, channel_chunk=64)
    np.testing.assert_allclose(host(f1), host(f2), rtol=1e-5, atol=1e-6)
    f3, _, _ = rasterization(**{**kw, "colors": feats[:, :3].contiguous()})
    np.testing.assert_allclose(host(f1[..., :3]), host(f3), rtol=1e-5, atol=1e-6)
    # antialiased mode scales opacities by the compensation factor
    aa, _, meta_aa = rasterization(**kw, rasterize_mode="antialiased")
    assert aa.shape == rgb.shape
    nrm, _, _ = rasterization(**kw, render_mode="Normal")
    assert nrm.shape == (2, Ht, Wd, 3)
    with pytest.raises(ValueError):
        W.rasterize_to_pixels(meta["means2d"], meta["conics"], torch.rand(2, N, 514, device=device),
                              meta["opacities"], meta["betas"], Wd, Ht, 16, meta["isect_offsets"],
                              meta["flatten_ids"])


def test_quats_path_autograd(device):
    """rasterization() without covars reads l_triagnles as quaternions."""
    from util import opaque_scene

    s = opaque_scene(2000, W=128, H=128, seed=3)
    t = {k: torch.tensor(v, device=device) for k, v in s.items()}
    means = t["means"].requires_grad_(True)
    quats = t["quats"].requires_grad_(True)
    scales = t["scales"].requires_grad_(True)
    img, a, meta = rasterization(means, quats, scales, t["opac"], t["beta"], t["rgb"], t["viewmats"], t["Ks"], 128,
                                 128)
    img.sum().backward()
    for g in (means.grad, quats.grad, scales.grad):
        assert torch.isfinite(g).all() and g.abs().sum() > 0


def test_backward_fp64_transmittance_agrees(device):
    N, Wd, Ht = 4000, 128, 128
    sc = scenes.make_scene(N, 6, seed=6)
    vms, Ks, ts = scenes.make_cameras(1, Wd, Ht, seed=6)
    grads = []
    for f64 in (False, True):
        W.set_backward_fp64_transmittance(f64)
        p = beta.BetaParams.from_scene(sc, device)
        img, _, _ = beta.render(p, torch.tensor(vms[0], device=device), torch.tensor(Ks[0], device=device), Wd, Ht)
        img.square().sum().backward()
        grads.append(host(p.xyz.grad))
    W.set_backward_fp64_transmittance(False)
    assert rel_l2(grads[0], grads[1]) < 1e-4


def test_atomic_backward_is_repeatable(device):
    """The blend backward adds per-(tile, primitive) totals with hardware
    float atomics (the reference uses gpuAtomicAdd, rasterize_to_pixels_bwd.cu:
    255-272), so the order of the adds varies between runs: two backward
    passes over the same inputs must agree to rounding -- every leaf gradient
    within a relative L2 of 1e-6 and a max-abs of 1e-5 of its largest entry."""
    N, Wd, Ht = 200_000, 640, 480
    sc = scenes.make_scene(N, 6, seed=11)
    vms, Ks, _ = scenes.make_cameras(1, Wd, Ht, seed=11)
    p = beta.BetaParams.from_scene(sc, device)
    tgt = torch.rand(Ht, Wd, 3, generator=torch.Generator().manual_seed(3)).to(device)
    vm, K = torch.tensor(vms, device=device), torch.tensor(Ks, device=device)
    grads = []
    for _ in range(2):
        for t in p.leaves().values():
            t.grad = None
        img, _, meta = beta.render_views(p, vm, K, Wd, Ht)
        (img[0] - tgt).abs().mean().backward()
        grads.append({k: t.grad.clone() for k, t in p.leaves().items()})
    assert int(meta["isect_ids"].numel()) > 500_000
    for k in grads[0]:
        a, b = grads[0][k].double(), grads[1][k].double()
        scale = b.abs().max().item()
        assert scale > 0, k
        rel = ((a - b).norm() / b.norm()).item()
        assert rel <= 1e-6, (k, rel)
        assert (a - b).abs().max().item() <= 1e-5 * scale, (k, (a - b).abs().max().item(), scale)
