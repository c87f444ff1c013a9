"""The hand-written backward behind the autograd Functions, checked by
central finite differences: a 7-D dynamic scene at 128x128, 64 random
primitives among those whose gradient is within three decades of the
largest (a primitive behind T ~ 1e-4 of cover has a gradient that is all
rounding, in FP64 and fp32 alike), the GPU's fp32 gradient of the L1 loss
(fused projection path, exact conditioning gradient) against FP64 central
differences of the oracle's forward in the GPU's tile lists, along a random
direction in each primitive's leaves.

The loss is piecewise smooth: a perturbation can cross the T cut-off, a
support edge or the 0.999 opacity clamp, and the current point itself can
sit on such a kink.  Each direction is differenced centrally at eps = 1e-3
and 5e-4, and one-sidedly to second order, (-3 f(0) + 4 f(h) - f(2h)) / 2h
with h = 5e-4 on either side.  It counts as smooth when the two central
differences agree to 2e-2 and the two one-sided ones to 1e-2 (relative);
the Richardson value (4 d(eps/2) - d(eps)) / 3 is then the reference.  On
this scene the oracle's own FP64 gradient meets that reference to 1.2e-4 at
worst, 1.2e-7 at the median.  Tolerance for the GPU: at least 48 of the 64
directions smooth, and on those a relative error <= 1e-4 at the median,
<= 1e-3 for 90% and <= 1e-2 for all; every direction (smooth or not)
within 2e-2 of the eps = 1e-3 difference for at least 60 of 64, median
<= 5e-3.
"""
import numpy as np
import pytest
import torch

from gsplat import beta, scenes
from gsplat.cuda import _wrapper as W
from oracle import ubs_oracle as O

pytestmark = pytest.mark.gpu

LEAVES = ("xyz", "mean_rest", "scale_raw", "l_triangle", "opacity_raw", "beta_raw", "rgb", "quats")


def test_backward_central_differences_7d(device):
    sc = scenes.make_scene(3000, 7, seed=31)
    sc.scale_raw[:, :3] += 1.5  # primitives of a few tiles at 128x128
    Wd = Ht = 128
    vms, Ks, ts = scenes.make_cameras(1, Wd, Ht, seed=31, radius=2.5)
    t = float(ts[0])
    tgt = np.random.Generator(np.random.PCG64(3)).uniform(0, 1, (Ht, Wd, 3))
    p = beta.BetaParams.from_scene(sc, device)
    W.set_cond_backward_exact(True)
    try:
        img, _, meta = beta.render_views(p, torch.tensor(vms, device=device), torch.tensor(Ks, device=device), Wd, Ht,
                                         timestamps=torch.tensor([t], device=device))
        (img[0] - torch.tensor(tgt, dtype=torch.float32, device=device)).abs().mean().backward()
    finally:
        W.set_cond_backward_exact(False)
    g = {k: getattr(p, k).grad.detach().cpu().numpy().astype(np.float64) for k in LEAVES}
    struct = tuple(meta[k].cpu().numpy() for k in ("tiles_per_gauss", "isect_ids", "flatten_ids", "isect_offsets"))
    L = {k: getattr(sc, k).astype(np.float64) for k in LEAVES}
    w = np.sign(img[0].detach().cpu().numpy().astype(np.float64) - tgt) / (Ht * Wd * 3)  # d loss / d image
    _, _, om = O.beta_step(L, vms[0], Ks[0], Wd, Ht, tgt, timestamp=t, structure=struct, want_grads=False)
    q = om["query"]

    def surrogate(lv):
        im, _, _ = O.beta_step(lv, vms[0], Ks[0], Wd, Ht, tgt, timestamp=t, structure=struct, want_grads=False,
                               fixed_query=q)
        return (im * w).sum()

    mag = sum(np.abs(g[k]).sum(axis=1) for k in LEAVES)
    reached = np.nonzero(mag >= 1e-3 * mag.max())[0]
    assert reached.size >= 128, reached.size
    rng = np.random.Generator(np.random.PCG64(4))
    pick = rng.choice(reached, 64, replace=False)
    f0 = surrogate(L)
    rel, rel_smooth, worst = [], [], []
    for n in pick:
        u = {k: np.zeros_like(v) for k, v in L.items()}
        for k in LEAVES:
            u[k][n] = rng.normal(0, 1, u[k].shape[1:])
        an = sum((g[k] * u[k]).sum() for k in LEAVES)
        fp1, fm1 = surrogate({k: L[k] + 1e-3 * u[k] for k in L}), surrogate({k: L[k] - 1e-3 * u[k] for k in L})
        fp2, fm2 = surrogate({k: L[k] + 5e-4 * u[k] for k in L}), surrogate({k: L[k] - 5e-4 * u[k] for k in L})
        d1, d2 = (fp1 - fm1) / 2e-3, (fp2 - fm2) / 1e-3
        fwd, bwd = (4.0 * fp2 - fp1 - 3.0 * f0) / 1e-3, (3.0 * f0 - 4.0 * fm2 + fm1) / 1e-3
        rich = (4.0 * d2 - d1) / 3.0
        scale = max(abs(rich), 1e-30)
        rel.append(abs(an - d1) / max(abs(d1), 1e-30))
        smooth = abs(d1 - d2) <= 2e-2 * scale and abs(fwd - bwd) <= 1e-2 * scale
        if smooth:
            rel_smooth.append(abs(an - rich) / scale)
        worst.append((rel[-1], int(n), an, d1, rich, smooth))
    rel, rs = np.array(rel), np.array(rel_smooth)
    print("all 64: median %.2e, p90 %.2e, max %.2e; smooth %d: median %.2e, p90 %.2e, max %.2e" % (
        np.median(rel), np.percentile(rel, 90), rel.max(), rs.size, np.median(rs), np.percentile(rs, 90), rs.max()),
        flush=True)
    for r, n, an, d1, rich, smooth in sorted(worst)[-4:]:
        print(f"  primitive {n}: rel {r:.2e} analytic {an:.6e} d(1e-3) {d1:.6e} richardson {rich:.6e} smooth {smooth}",
              flush=True)
    assert (rel <= 2e-2).sum() >= 60, np.sort(rel)[-8:]
    assert np.median(rel) <= 5e-3, np.median(rel)
    assert rs.size >= 48, rs.size
    assert np.median(rs) <= 1e-4 and np.percentile(rs, 90) <= 1e-3 and rs.max() <= 1e-2, np.sort(rs)[-6:]
