"""The `_C` stand-in (gsplat/cuda/_c_shim.py) called the way the reference's
_wrapper.py calls its extension (argument order and return tuples of
cuda/csrc/bindings.h), against this package's public API on the same
inputs.  Same kernels underneath, so results are identical, except the
blend backward, which here re-derives T from alpha (the reference's
interface has no T_final) and is compared within tolerance."""
import numpy as np
import pytest
import torch

from gsplat import scenes
from gsplat.cuda import _c_shim as _C
from gsplat.cuda import _wrapper as W

from util import conditioned_scene, opaque_scene, rel_l2

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_shim_matches_public_api(device):
    # conditioning + covariance
    c = conditioned_scene(2000, 6, seed=4)
    a = c["act"]
    ins = [cu(a["mean"]), cu(c["cov"]), cu(a["opac"]), cu(np.ascontiguousarray(a["beta"][:, 1:])), cu(c["query"])]
    m1 = _C.cond_mean_convariance_opacity_fwd(*ins)
    m2 = W.cond_mean_convariance_opacity(*ins)
    for x, y in zip(m1, m2):
        assert torch.equal(x, y)
    g = [torch.randn_like(t) for t in m1]
    b1 = _C.cond_mean_convariance_opacity_bwd(*ins, *g)
    tins = [t.clone().requires_grad_(i < 4) for i, t in enumerate(ins)]
    torch.autograd.backward(list(W.cond_mean_convariance_opacity(*tins)), g)
    for x, t in zip(b1, tins[:4]):
        assert torch.equal(x, t.grad)

    # projection -> isect -> offsets -> blend, reference call order
    s = opaque_scene(3000, W=200, H=136, seed=5)
    means, cov6, vm, K = cu(s["means"]), cu(s["covars6"]), cu(s["viewmats"]), cu(s["Ks"])
    p1 = _C.fully_fused_projection_fwd(means, cov6, None, None, vm, K, 200, 136, 0.3, 0.01, 1e10, 0.0, False, False)
    p2 = W.fully_fused_projection(means, cov6, None, None, vm, K, 200, 136)
    vis = p2[0] > 0
    assert torch.equal(p1[0], p2[0])
    for x, y in zip(p1[1:4], p2[1:4]):
        assert torch.equal(x[vis], y[vis])
    radii, m2d, depths, conics = p2[:4]
    tw, th = (200 + 15) // 16, (136 + 15) // 16
    i1 = _C.isect_tiles(m2d, radii, depths, None, None, 1, 16, tw, th, True, True)
    i2 = W.isect_tiles(m2d, radii, depths, 16, tw, th)
    for x, y in zip(i1, i2):
        assert torch.equal(x, y)
    off = _C.isect_offset_encode(i1[1], 1, tw, th)
    assert torch.equal(off, W.isect_offset_encode(i2[1], 1, tw, th))
    op = cu(s["opac"])[None]
    be = cu(s["beta"])[None]
    col = cu(s["rgb"])[None]
    r1 = _C.rasterize_to_pixels_fwd(m2d, conics, col, op, be, None, None, 200, 136, 16, off, i1[2])
    r2 = W.raster_fwd_raw(m2d, conics, col, op, be, None, None, 200, 136, 16, off, i1[2])
    for x, y in zip(r1, r2[:3]):
        assert torch.equal(x, y)
    vrc = torch.randn_like(r1[0])
    vra = torch.randn_like(r1[1])
    g1 = _C.rasterize_to_pixels_bwd(m2d, conics, col, op, be, None, None, 200, 136, 16, off, i1[2], r1[1], r1[2],
                                    vrc, vra)
    g2 = W.raster_bwd_raw(m2d, conics, col, op, be, None, None, 200, 136, 16, off, i1[2], r2[1], r2[2], r2[3],
                          vrc, vra)
    for x, y, n in zip(g1, g2, ("v_means2d", "v_conics", "v_colors", "v_opac", "v_betas")):
        assert rel_l2(x.cpu().numpy(), y.cpu().numpy().astype(np.float64)) < 1e-2, n


def test_shim_rejects_packed(device):
    z = torch.zeros(1, 4, 2, device="cuda")
    with pytest.raises(NotImplementedError):
        _C.isect_tiles(z, torch.zeros(1, 4, dtype=torch.int32, device="cuda"), torch.zeros(1, 4, device="cuda"),
                       torch.zeros(4, dtype=torch.int64, device="cuda"), None, 1, 16, 4, 4, True, True)
