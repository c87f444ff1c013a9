"""Stored-transmittance blend pair (ubs_rasterize_to_pixels_fwd_ckpt /
_bwd_ckpt, gsplat.cuda._wrapper.set_backward_transmittance_checkpoint).

* The forward's outputs are those of the plain forward, bit for bit.
* Each stored value is the pixel's T after its batch: checked against the
  FP64 oracle blending the tile lists cut after that batch (to 2e-5, on
  pixels away from the T cut-off), and the batch holding the pixel's last
  contributor stores T_final exactly.
* The backward restarting every batch from the stored T matches the FP64
  oracle (relative L2 <= 2e-3, the tolerance of the recomputing backward)
  and the recomputing backward (relative L2 <= 1e-3), over tile sizes,
  channel chunks, cameras and tile masks, on lists several batches long.
"""
import numpy as np
import pytest
import torch

from gsplat import beta, scenes
from gsplat.cuda import _wrapper as W
from oracle import ubs_oracle as O

from util import assert_same, check_branch_pixels, opaque_scene, rel_l2

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.tensor(np.ascontiguousarray(a), device="cuda")


def host(t):
    return t.detach().cpu().numpy()


def _inputs(N, W_, H_, C, ts, channels, opacity):
    s = opaque_scene(N, W=W_, H=H_, n_cams=C, opacity=opacity, scale=(0.03, 0.15))
    r, m2, dp, cn, _ = O.proj_fwd(s["means"], s["covars6"], None, None, s["viewmats"], s["Ks"], W_, H_)
    tw, th = (W_ + ts - 1) // ts, (H_ + ts - 1) // ts
    _, ids, flat = O.isect_tiles(m2, r, dp, ts, tw, th)
    off = O.isect_offset_encode(ids, C, tw, th)
    g = np.random.Generator(np.random.PCG64(channels))
    rgb = s["rgb"] if channels == 3 else g.uniform(0, 1, (N, channels)).astype(np.float32)
    return dict(m2=m2, cn=cn, col=np.ascontiguousarray(np.broadcast_to(rgb, (C,) + rgb.shape)),
                op=np.ascontiguousarray(np.broadcast_to(s["opac"], (C, N))),
                be=np.ascontiguousarray(np.broadcast_to(s["beta"], (C, N))), off=off, flat=flat, tw=tw, th=th)


def _ckpt_tile(ck, slot, ts):
    """The TS x TS pixels of one checkpoint slot, in image order (thread t
    holds pixel (sy*8 + lane//8, sx*8 + lane%8) of wave t//64 = sy*SB + sx)."""
    BS, SB = ts * ts, ts // 8
    out = np.zeros((ts, ts), np.float32)
    vals = ck[slot * BS:(slot + 1) * BS]
    for t in range(BS):
        wave, lane = t >> 6, t & 63
        sx, sy = wave % SB, wave // SB
        out[sy * 8 + (lane >> 3), sx * 8 + (lane & 7)] = vals[t]
    return out


def _truncated(x, C, k, BS):
    """Tile lists cut after batch k (entries [s, s + (k+1) BS))."""
    off = x["off"].reshape(-1)
    n = x["flat"].size
    ends = np.append(off[1:], n)
    new_flat, new_off, pos = [], [], 0
    for L in range(off.size):
        s, e = int(off[L]), int(ends[L])
        e2 = min(e, s + (k + 1) * BS)
        new_off.append(pos)
        new_flat.append(x["flat"][s:e2])
        pos += e2 - s
    return (np.asarray(new_off, np.int32).reshape(x["off"].shape),
            np.concatenate(new_flat).astype(np.int32) if new_flat else np.zeros(0, np.int32))


@pytest.mark.parametrize("C,channels,ts,mask", [
    (1, 3, 16, False), (2, 3, 8, False), (1, 4, 32, False), (1, 35, 8, True), (2, 7, 16, True)])
def test_checkpoint_pair(device, C, channels, ts, mask):
    W_, H_ = 128, 96
    x = _inputs(6000, W_, H_, C, ts, channels, opacity=(0.03, 0.3))
    BS = ts * ts
    off_flat = x["off"].reshape(-1)
    lens = np.diff(np.append(off_flat, x["flat"].size))
    assert lens.max() > 2 * BS  # several batches per tile
    g = np.random.Generator(np.random.PCG64(7 + channels))
    masks = (g.uniform(0, 1, (C, x["th"], x["tw"])) > 0.25) if mask else None
    tmask = None if masks is None else cu(masks.astype(np.bool_))
    ins = [cu(x[k]) for k in ("m2", "cn", "col", "op", "be")]
    plain = W.raster_fwd_raw(*ins, None, tmask, W_, H_, ts, cu(x["off"]), cu(x["flat"]))
    rc, ra, last, Tf, ck = W.raster_fwd_ckpt_raw(*ins, None, tmask, W_, H_, ts, cu(x["off"]), cu(x["flat"]))
    for a, b, n in zip((rc, ra, last, Tf), plain, ("colors", "alphas", "last_ids", "T_final")):
        assert_same(host(a), host(b), n)
    assert ck.numel() == (x["flat"].size // BS + C * x["th"] * x["tw"] + 1) * BS
    d = np.float64
    want_c, want_a, want_last, mT, mS = O.raster_fwd(*(x[k].astype(d) for k in ("m2", "cn", "col", "op", "be")),
                                                    W_, H_, ts, x["off"], x["flat"], masks=masks, margins=True)
    safe = (mT > 1e-3) & (mS > 1e-6)
    if masks is not None:
        off_px = ~np.repeat(np.repeat(masks, ts, axis=1), ts, axis=2)[:, :H_, :W_]
        safe |= off_px
    assert safe.mean() > 0.9
    c32, _, last32 = O.raster_fwd(x["m2"], x["cn"], x["col"], x["op"], x["be"], W_, H_, ts, x["off"], x["flat"],
                                  masks=masks)
    check_branch_pixels(host(rc), host(last), c32, last32, np.ones(safe.shape, bool), safe,
                        what=f"checkpointed blend fwd channels={channels}")
    # stored values: batch 0 and 1 against the FP64 blend of the cut lists;
    # the batch of the pixel's last contributor stores T_final exactly
    ckh, lasth, Tfh = host(ck), host(last), host(Tf)
    checked = 0
    for k in (0, 1):
        toff, tflat = _truncated(x, C, k, BS)
        _, ta, _ = O.raster_fwd(*(x[q].astype(d) for q in ("m2", "cn", "col", "op", "be")), W_, H_, ts, toff,
                                tflat, masks=masks)
        for L in range(off_flat.size):
            cam, tile = divmod(L, x["th"] * x["tw"])
            ty, tx = divmod(tile, x["tw"])
            if masks is not None and not masks[cam, ty, tx]:
                continue
            s = int(off_flat[L])
            if lens[L] <= k * BS:
                continue
            ys, xs = slice(ty * ts, min(H_, ty * ts + ts)), slice(tx * ts, min(W_, tx * ts + ts))
            lt = lasth[cam, ys, xs]
            if lt.max() < s + k * BS:  # the forward may have stopped before this batch
                continue
            got = _ckpt_tile(ckh, s // BS + L + k, ts)[:lt.shape[0], :lt.shape[1]]
            want = 1.0 - ta[cam, ys, xs, 0]
            ok = safe[cam, ys, xs]
            assert np.abs(got - want)[ok].max(initial=0.0) <= 2e-5, (k, L)
            in_batch = (lt >= s + k * BS) & (lt < s + (k + 1) * BS) & (lt >= s)
            fin = in_batch & (lt == lt)  # pixels whose last contributor is in this batch
            if (k + 1) * BS >= lens[L]:  # the tile's last batch: every pixel stores T_final
                fin = np.ones_like(in_batch)
            np.testing.assert_array_equal(got[fin], Tfh[cam, ys, xs][fin])
            checked += int(ok.sum())
    assert checked > 1000
    vrc = (g.normal(0, 1, host(rc).shape) * safe[..., None]).astype(np.float32)
    vra = (g.normal(0, 1, host(ra).shape) * safe[..., None]).astype(np.float32)
    got = W.raster_bwd_raw(*ins, None, tmask, W_, H_, ts, cu(x["off"]), cu(x["flat"]), ra, last, Tf, cu(vrc),
                           cu(vra), T_ckpt=ck)
    rec = W.raster_bwd_raw(*ins, None, tmask, W_, H_, ts, cu(x["off"]), cu(x["flat"]), ra, last, Tf, cu(vrc),
                           cu(vra))
    want = O.raster_bwd(*(x[k].astype(d) for k in ("m2", "cn", "col", "op", "be")), W_, H_, ts, x["off"],
                        x["flat"], want_a, want_last, vrc.astype(d), vra.astype(d), masks=masks,
                        T_final=1.0 - want_a[..., 0])
    for a, b, r, n in zip(got, want, rec, ("v_means2d", "v_conics", "v_colors", "v_opac", "v_betas")):
        assert np.abs(b).max() > 0, n
        assert rel_l2(host(a), b) < 2e-3, (n, rel_l2(host(a), b))
        assert rel_l2(host(a), host(r)) < 1e-3, (n, rel_l2(host(a), host(r)))


def test_checkpoint_switch_in_autograd(device):
    sc = scenes.make_scene(20_000, 6, seed=3)
    vms, Ks, _ = scenes.make_cameras(2, 160, 128, seed=3)
    tgt = torch.rand((2, 128, 160, 3), generator=torch.Generator().manual_seed(1)).to(device)
    out = {}
    for mode in (False, True):
        p = beta.BetaParams.from_scene(sc, device)
        W.set_backward_transmittance_checkpoint(mode)
        try:
            img, _, _ = beta.render_views(p, torch.tensor(vms, device=device), torch.tensor(Ks, device=device),
                                          160, 128)
            (img - tgt).abs().mean().backward()
        finally:
            W.set_backward_transmittance_checkpoint(False)
        out[mode] = (img.detach(), {k: t.grad.detach().clone() for k, t in p.leaves().items()})
    assert torch.equal(out[False][0], out[True][0])
    for k, g in out[False][1].items():
        assert rel_l2(host(out[True][1][k]), host(g)) < 1e-3, k
    with torch.no_grad():  # no gradient wanted: no checkpoints are stored
        W.set_backward_transmittance_checkpoint(True)
        try:
            p = beta.BetaParams.from_scene(sc, device, requires_grad=False)
            img2, _, _ = beta.render_views(p, torch.tensor(vms, device=device), torch.tensor(Ks, device=device),
                                           160, 128)
        finally:
            W.set_backward_transmittance_checkpoint(False)
    assert torch.equal(img2, out[False][0])
