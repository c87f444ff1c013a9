"""Pins the oracle's tile intersection, sort, ranges and blending against
brute force, numpy primitives, closed forms and finite differences.  CPU only."""
import math

import numpy as np
import pytest

from gsplat import scenes
from oracle import ubs_oracle as O


# ---------------------------------------------------------------- intersection
def brute_isect(means2d, radii, depths, ts, tw, th):
    """Python loops over every tile of every primitive (float32 box math)."""
    C, N = radii.shape
    tbits = int(tw * th).bit_length()
    keys, flat, tpg = [], [], np.zeros((C, N), np.int32)
    f = np.float32
    for c in range(C):
        for n in range(N):
            r = radii[c, n]
            if r <= 0:
                continue
            tr = f(r) / f(ts)
            tx, ty = f(means2d[c, n, 0]) / f(ts), f(means2d[c, n, 1]) / f(ts)
            sat = lambda v: 0 if not v > 0 else int(v)  # CUDA float->uint saturation
            x0, y0 = min(sat(math.floor(tx - tr)), tw), min(sat(math.floor(ty - tr)), th)
            x1, y1 = min(sat(math.ceil(tx + tr)), tw), min(sat(math.ceil(ty + tr)), th)
            tpg[c, n] = (y1 - y0) * (x1 - x0)
            dbits = int(np.array(depths[c, n], np.float32).view(np.int32))
            for i in range(y0, y1):
                for j in range(x0, x1):
                    keys.append((c << (32 + tbits)) | ((i * tw + j) << 32) | (dbits & 0xFFFFFFFF))
                    flat.append(c * N + n)
    return tpg, np.array(keys, np.int64), np.array(flat, np.int32)


@pytest.mark.parametrize("C", [1, 3])
def test_isect_brute_force_and_stable_sort(C):
    g = np.random.Generator(np.random.PCG64(C))
    N, W, H, ts = 300, 100, 70, 16
    tw, th = math.ceil(W / ts), math.ceil(H / ts)
    means2d = g.uniform(-30, 130, (C, N, 2)).astype(np.float32)
    radii = g.integers(-2, 25, (C, N)).astype(np.int32)
    depths = g.choice(np.float32([0.5, 1.0, 2.5, 3.0]), (C, N))  # ties: order decided by stability
    tpg, keys, flat = brute_isect(means2d, radii, depths, ts, tw, th)
    got_tpg, got_u, got_uf = O.isect_tiles(means2d, radii, depths, ts, tw, th, sort=False)
    np.testing.assert_array_equal(got_tpg, tpg)
    np.testing.assert_array_equal(got_u, keys)
    np.testing.assert_array_equal(got_uf, flat)
    _, got_s, got_sf = O.isect_tiles(means2d, radii, depths, ts, tw, th, sort=True)
    end_bit = 32 + int(tw * th).bit_length() + int(C).bit_length()
    order = np.argsort(keys & ((1 << end_bit) - 1), kind="stable")
    np.testing.assert_array_equal(got_s, keys[order])
    np.testing.assert_array_equal(got_sf, flat[order])
    # ranges: first index of each (camera, tile) = searchsorted of the prefix
    off = O.isect_offset_encode(got_s, C, tw, th)
    prefix = (got_s >> 32)
    ids = (prefix >> int(tw * th).bit_length()) * (tw * th) + (prefix & ((1 << int(tw * th).bit_length()) - 1))
    np.testing.assert_array_equal(off.reshape(-1), np.searchsorted(ids, np.arange(C * tw * th), side="left"))


def test_sort_pairs_bits_and_stability():
    g = np.random.Generator(np.random.PCG64(9))
    keys = g.integers(-(1 << 63), (1 << 63) - 1, 5000, dtype=np.int64)
    vals = np.arange(5000, dtype=np.int32)
    for lo, hi in ((0, 64), (0, 13), (8, 40), (20, 21)):
        k, v = O.sort_pairs(keys, vals, lo, hi)
        m = np.uint64(((1 << hi) - 1) & ~((1 << lo) - 1))
        order = np.argsort(keys.view(np.uint64) & m, kind="stable")
        np.testing.assert_array_equal(v, vals[order])
        np.testing.assert_array_equal(k, keys[order])


def test_offsets_empty_and_edges():
    off = O.isect_offset_encode(np.zeros(0, np.int64), 2, 3, 2)
    assert (off == 0).all()
    ids = np.array([(0 << 35) | (5 << 32), (1 << 35) | (0 << 32), (1 << 35) | (0 << 32)], np.int64)  # tbits=3
    off = O.isect_offset_encode(ids, 2, 3, 2)
    np.testing.assert_array_equal(off.reshape(-1), [0, 0, 0, 0, 0, 0, 1, 3, 3, 3, 3, 3])


# ---------------------------------------------------------------- blending
def one_prim(x, y, conic, o, b, color, W=16, H=16, bg=None):
    means2d = np.array([[[x, y]]], np.float64)
    conics = np.array([[conic]], np.float64)
    colors = np.array([[color]], np.float64)
    op, be = np.array([[o]], np.float64), np.array([[b]], np.float64)
    off = np.zeros((1, 1, 1), np.int32)
    flat = np.zeros(1, np.int32)
    return O.raster_fwd(means2d, conics, colors, op, be, W, H, 16, off, flat, bg)


def test_single_primitive_closed_form():
    x, y, conic, o, b = 7.3, 8.1, [0.02, 0.005, 0.03], 0.8, 2.5
    col = [0.2, 0.5, 0.9]
    bg = np.array([[0.1, 0.2, 0.3]])
    c, a, last = one_prim(x, y, conic, o, b, col, bg=bg)
    for i in range(16):
        for j in range(16):
            dx, dy = x - (j + 0.5), y - (i + 0.5)
            s = conic[0] * dx * dx + conic[2] * dy * dy + 2 * conic[1] * dx * dy
            al = min(0.999, o * (1 - s) ** b) if 0 <= s < 1 else 0.0
            np.testing.assert_allclose(a[0, i, j, 0], al, rtol=1e-12, atol=1e-15)
            np.testing.assert_allclose(c[0, i, j], np.array(col) * al + (1 - al) * bg[0], rtol=1e-12, atol=1e-15)
            assert last[0, i, j] == 0


def test_two_primitives_and_cutoff():
    # two coincident flat primitives: C = c1 a1 + c2 a2 (1 - a1)
    means2d = np.array([[[8.0, 8.0], [8.0, 8.0]]])
    conics = np.array([[[1e-4, 0, 1e-4], [1e-4, 0, 1e-4]]])
    colors = np.array([[[1.0, 0, 0], [0, 1.0, 0]]])
    be = np.array([[1.0, 1.0]])
    off = np.zeros((1, 1, 1), np.int32)
    flat = np.array([0, 1], np.int32)
    for o1, o2, expect_second in ((0.5, 0.7, True), (0.999, 0.95, False)):
        op = np.array([[o1, o2]])
        c, a, last = O.raster_fwd(means2d, conics, colors, op, be, 16, 16, 16, off, flat)
        s = 1e-4 * (0.5 ** 2 + 0.5 ** 2)
        a1, a2 = min(0.999, o1 * (1 - s)), min(0.999, o2 * (1 - s))
        # pixel (8,8): both at distance (0.5, 0.5)
        if expect_second:
            np.testing.assert_allclose(c[0, 8, 8], [a1, a2 * (1 - a1), 0], rtol=1e-12)
            np.testing.assert_allclose(a[0, 8, 8, 0], 1 - (1 - a1) * (1 - a2), rtol=1e-12)
            assert last[0, 8, 8] == 1
        else:
            # (1 - 0.999)(1 - a2) <= 1e-4: the second primitive is excluded
            assert (1 - a1) * (1 - a2) <= 1e-4
            np.testing.assert_allclose(c[0, 8, 8], [a1, 0, 0], rtol=1e-12)
            assert last[0, 8, 8] == 0


def test_sigma_domain_and_empty_tile():
    # sigma < 0 (indefinite conic) and sigma >= 1 contribute nothing
    c, a, _ = one_prim(8.0, 8.0, [-1.0, 0.0, -1.0], 0.9, 2.0, [1, 1, 1])
    assert (a == 0).all()
    c, a, _ = one_prim(100.0, 100.0, [0.01, 0.0, 0.01], 0.9, 2.0, [1, 1, 1])
    assert (a == 0).all()
    # empty tiles render the background
    off = np.zeros((1, 2, 2), np.int32)
    bg = np.array([[0.3, 0.4, 0.5]])
    c, a, _ = O.raster_fwd(np.zeros((1, 1, 2)), np.zeros((1, 1, 3)), np.zeros((1, 1, 3)), np.zeros((1, 1)),
                           np.ones((1, 1)), 32, 32, 16, off, np.zeros(0, np.int32), bg)
    np.testing.assert_array_equal(c[0], np.broadcast_to(bg[0], (32, 32, 3)))


def small_scene(seed=0, N=60, W=48, H=40, opacity=(0.2, 0.8)):
    g = np.random.Generator(np.random.PCG64(seed))
    means2d = g.uniform(0, [W, H], (1, N, 2))
    A = g.normal(0, 1, (N, 2, 2))
    cov = A @ A.transpose(0, 2, 1) * g.uniform(4, 40, (N, 1, 1)) + 2 * np.eye(2)
    inv = np.linalg.inv(cov)
    conics = np.stack([inv[:, 0, 0], inv[:, 0, 1], inv[:, 1, 1]], -1)[None]
    colors = g.uniform(0, 1, (1, N, 3))
    op = g.uniform(*opacity, (1, N))
    be = g.uniform(1.5, 5.0, (1, N))
    depths = g.uniform(1, 5, (1, N)).astype(np.float32)
    radii = np.full((1, N), 30, np.int32)
    tw, th = math.ceil(W / 16), math.ceil(H / 16)
    _, ids, flat = O.isect_tiles(means2d.astype(np.float32), radii, depths, 16, tw, th)
    off = O.isect_offset_encode(ids, 1, tw, th)
    return dict(means2d=means2d, conics=conics, colors=colors, op=op, be=be, off=off, flat=flat, W=W, H=H)


def test_raster_bwd_central_differences():
    s = small_scene(3)
    args = lambda m2, cn, col, op, be: (m2, cn, col, op, be, s["W"], s["H"], 16, s["off"], s["flat"])
    c, a, last, mT, mS = O.raster_fwd(*args(s["means2d"], s["conics"], s["colors"], s["op"], s["be"]), margins=True)
    assert (mT > 1e-2).all() and (mS > 1e-4).all()
    g = np.random.Generator(np.random.PCG64(4))
    vc, va = g.normal(0, 1, c.shape), g.normal(0, 1, a.shape)
    grads = O.raster_bwd(*args(s["means2d"], s["conics"], s["colors"], s["op"], s["be"]), a, last, vc, va)
    names = ["means2d", "conics", "colors", "op", "be"]
    for k, name in enumerate(names):
        base = [s[n] for n in names]
        d = g.normal(0, 1, base[k].shape)
        eps = 1e-6 * (1e-3 if name == "conics" else 1.0)

        def f(x):
            b2 = list(base)
            b2[k] = x
            cc, aa, _ = O.raster_fwd(*args(*b2))
            return (cc * vc).sum() + (aa * va).sum()

        fd = (f(base[k] + eps * d) - f(base[k] - eps * d)) / (2 * eps)
        np.testing.assert_allclose((grads[k] * d).sum(), fd, rtol=2e-5, err_msg=name)


def test_beta_step_central_differences():
    """Full step (activations -> conditioning -> projection -> blend -> L1)
    in FP64: exact-mode leaf gradients match central differences."""
    sc = scenes.make_scene(200, 6, seed=11)
    W, H = 64, 48
    vms, Ks, ts = scenes.make_cameras(1, W, H, seed=11, radius=2.5)
    leaves = {k: getattr(sc, k).astype(np.float64) for k in
              ("xyz", "mean_rest", "scale_raw", "l_triangle", "opacity_raw", "beta_raw", "rgb", "quats")}
    leaves["scale_raw"][:, :3] += 2.0  # larger primitives so many pixels are covered
    tgt = np.random.Generator(np.random.PCG64(1)).uniform(0, 1, (H, W, 3))
    img, grads, meta = O.beta_step(leaves, vms[0], Ks[0], W, H, tgt, exact_cond=True)
    struct = (meta["tiles_per_gauss"], meta["isect_ids"], meta["flatten_ids"], meta["isect_offsets"])
    # a smooth surrogate of the loss for FD: sum(img * w) with w = d loss / d img
    w = np.sign(img - tgt) / (H * W * 3)
    g = np.random.Generator(np.random.PCG64(2))
    for name in ("xyz", "mean_rest", "scale_raw", "l_triangle", "opacity_raw", "beta_raw", "rgb", "quats"):
        d = g.normal(0, 1, leaves[name].shape)
        eps = 1e-6

        def f(x):
            lv = dict(leaves)
            lv[name] = x
            im, _, _ = O.beta_step(lv, vms[0], Ks[0], W, H, tgt, structure=struct, want_grads=False,
                                   fixed_query=meta["query"])
            return (im * w).sum()

        fd = (f(leaves[name] + eps * d) - f(leaves[name] - eps * d)) / (2 * eps)
        an = (grads[name] * d).sum()
        assert abs(an - fd) <= 1e-4 * max(abs(fd), 1e-3 * np.abs(w).sum()), (name, an, fd)


def test_beta_step_tile_mask():
    """A tile mask restricts the blend and the loss to its tiles: an all-true
    mask changes nothing; a partial mask reproduces the unmasked image on its
    tiles, the background elsewhere, and the gradient of the loss taken over
    its pixels only (checked against the unmasked run with that loss)."""
    sc = scenes.make_scene(300, 6, seed=12)
    W, H = 64, 48
    vms, Ks, _ = scenes.make_cameras(1, W, H, seed=12, radius=2.5)
    leaves = {k: getattr(sc, k).astype(np.float64) for k in
              ("xyz", "mean_rest", "scale_raw", "l_triangle", "opacity_raw", "beta_raw", "rgb", "quats")}
    leaves["scale_raw"][:, :3] += 2.0
    tgt = np.random.Generator(np.random.PCG64(3)).uniform(0, 1, (H, W, 3))
    img, grads, _ = O.beta_step(leaves, vms[0], Ks[0], W, H, tgt)
    full = np.ones((3, 4), bool)
    img_f, grads_f, _ = O.beta_step(leaves, vms[0], Ks[0], W, H, tgt, tile_mask=full)
    np.testing.assert_array_equal(img_f, img)
    for k in grads:
        np.testing.assert_array_equal(grads_f[k], grads[k])
    tm = np.zeros((3, 4), bool)
    tm[0, 1] = tm[2, 3] = tm[1, 0] = True
    pm = np.kron(tm, np.ones((16, 16), bool))
    img_m, grads_m, _ = O.beta_step(leaves, vms[0], Ks[0], W, H, tgt, tile_mask=tm)
    np.testing.assert_array_equal(img_m[pm], img[pm])
    assert np.all(img_m[~pm] == 0.0)  # black background
    # unmasked run with the loss over the masked pixels: the target equals the image elsewhere
    tgt2 = np.where(pm[..., None], tgt, img)
    _, grads2, _ = O.beta_step(leaves, vms[0], Ks[0], W, H, tgt2)
    scale = (H * W) / pm.sum()  # mean over all pixels vs over the masked ones
    for k in grads:
        np.testing.assert_allclose(grads_m[k], grads2[k] * scale, rtol=1e-9, atol=1e-15)
