"""GPU parity of every kernel against the oracle (same seeded inputs).

Bit-exact: rotmat, covariance forward, conditioned mean/covariance, projection
(radii, means2d, depths, conics), tiles_per_gauss, isect ids, flatten ids,
sort order, tile offsets.  Tolerances (stated per test) for transcendental and
atomically accumulated outputs.
"""
import numpy as np
import pytest
import torch

from gsplat import scenes
from gsplat.cuda import _wrapper as W
from oracle import ubs_oracle as O

from util import assert_same, check_branch_pixels, conditioned_scene, opaque_scene, rel_l2

pytestmark = pytest.mark.gpu


def cu(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t if dtype is None else t.to(dtype)


def host(t):
    return t.detach().cpu().numpy()


# ------------------------------------------------------------------ covariance
def test_rotmat_fwd_bwd_exact(device):
    g = np.random.Generator(np.random.PCG64(1))
    l = g.normal(0, 1, (5000, 3)).astype(np.float32)
    vR = g.normal(0, 1, (5000, 3, 3)).astype(np.float32)
    lt = cu(l).requires_grad_(True)
    R = W.l_triangle_to_rotmat(lt)
    assert_same(host(R), O.rotmat_fwd(l), "rotmat")
    R.backward(cu(vR))
    assert_same(host(lt.grad), O.rotmat_bwd(vR), "rotmat_bwd")


@pytest.mark.parametrize("D,spatial", [(6, False), (7, False), (6, True), (3, False), (9, False)])
def test_covar_fwd_exact_bwd_close(device, D, spatial):
    g = np.random.Generator(np.random.PCG64(D))
    N = 3000
    rot = scenes.quat_to_rotmat_np(g.normal(0, 1, (N, 4))).astype(np.float32)
    s = np.exp(g.normal(-1, 0.5, (N, D))).astype(np.float32)
    lt = g.normal(0, 0.2, (N, D * (D - 1) // 2)).astype(np.float32)
    ri, rj = scenes.rest_pattern(D)
    want = O.covar_fwd(rot, s, lt, ri, rj, spatial)
    tr, ts_, tl = cu(rot).requires_grad_(True), cu(s).requires_grad_(True), cu(lt).requires_grad_(True)
    got = W.rot_scale_l_triangle_to_covar(tr, ts_, tl, cu(ri), cu(rj), spatial)
    assert_same(host(got), want, f"covar D={D}")
    G = g.normal(0, 1, want.shape).astype(np.float32)
    got.backward(cu(G))
    wr, ws, wl = O.covar_bwd(rot.astype(np.float64), s.astype(np.float64), lt.astype(np.float64), ri, rj,
                             spatial, G.astype(np.float64))
    for a, b, n in ((tr.grad, wr, "v_rot"), (ts_.grad, ws, "v_scale"), (tl.grad, wl, "v_ltri")):
        assert rel_l2(host(a), b) < 1e-5, (n, rel_l2(host(a), b))


# ---------------------------------------------------------------- conditioning
@pytest.mark.parametrize("D", [6, 7])
def test_cond_fwd_exact_bwd_close(device, D):
    c = conditioned_scene(4000, D, seed=3)
    a = c["act"]
    ins = [a["mean"], c["cov"], a["opac"], np.ascontiguousarray(a["beta"][:, 1:]), c["query"]]
    m, V, o = O.cond_fwd(*ins)
    tins = [cu(x).requires_grad_(i < 4) for i, x in enumerate(ins)]
    gm, gV, go = W.cond_mean_convariance_opacity(*tins)
    assert_same(host(gm), m, "m_cond")
    assert_same(host(gV), V, "V_cond")
    # tanh / pow differ between the GPU and libm by a few ulp
    np.testing.assert_allclose(host(go), o, rtol=2e-5, atol=1e-30)
    g = np.random.Generator(np.random.PCG64(11))
    vm = g.normal(0, 1, m.shape).astype(np.float32)
    vV = g.normal(0, 1, V.shape).astype(np.float32)
    vo = g.normal(0, 1, o.shape).astype(np.float32)
    torch.autograd.backward([gm, gV, go], [cu(vm), cu(vV), cu(vo)])
    want = O.cond_bwd(*[x.astype(np.float64) for x in ins], vm.astype(np.float64), vV.astype(np.float64),
                      vo.astype(np.float64))
    for t, w, n in zip(tins[:4], want, ("g_means", "g_covars", "g_opac", "g_betas")):
        assert rel_l2(host(t.grad), w) < 2e-4, (n, rel_l2(host(t.grad), w))


@pytest.mark.parametrize("exact", [False, True])
def test_cond_bwd_modes(device, exact):
    """Both conditioning-backward modes against the oracle's same mode; the
    two modes differ only in the covariance gradient."""
    c = conditioned_scene(3000, 6, seed=8)
    a = c["act"]
    ins = [a["mean"], c["cov"], a["opac"], np.ascontiguousarray(a["beta"][:, 1:]), c["query"]]
    g = np.random.Generator(np.random.PCG64(12))
    vm, vV, vo = (g.normal(0, 1, s).astype(np.float32) for s in ((3000, 3), (3000, 3, 3), (3000, 1)))
    W.set_cond_backward_exact(exact)
    try:
        tins = [cu(x).requires_grad_(i < 4) for i, x in enumerate(ins)]
        outs = W.cond_mean_convariance_opacity(*tins)
        torch.autograd.backward(list(outs), [cu(vm), cu(vV), cu(vo)])
    finally:
        W.set_cond_backward_exact(False)
    d = np.float64
    want = O.cond_bwd(*[x.astype(d) for x in ins], vm.astype(d), vV.astype(d), vo.astype(d), exact=exact)
    other = O.cond_bwd(*[x.astype(d) for x in ins], vm.astype(d), vV.astype(d), vo.astype(d), exact=not exact)
    for t, w, n in zip(tins[:4], want, ("g_means", "g_covars", "g_opac", "g_betas")):
        assert rel_l2(host(t.grad), w) < 2e-4, (n, rel_l2(host(t.grad), w))
    assert rel_l2(host(tins[1].grad), other[1]) > 10 * rel_l2(host(tins[1].grad), want[1])


# ------------------------------------------------------------------ projection
@pytest.mark.parametrize("mode,C", [("covars", 1), ("covars", 3), ("quats", 1), ("quats", 2)])
def test_projection_fwd_exact(device, mode, C):
    s = opaque_scene(20000, W=320, H=200, n_cams=C, seed=C)
    cov = s["covars6"] if mode == "covars" else None
    q = s["quats"] if mode == "quats" else None
    sc = s["scales"] if mode == "quats" else None
    want = O.proj_fwd(s["means"], cov, q, sc, s["viewmats"], s["Ks"], 320, 200, calc_compensations=True)
    got = W.fully_fused_projection(cu(s["means"]), None if cov is None else cu(cov), None if q is None else cu(q),
                                   None if sc is None else cu(sc), cu(s["viewmats"]), cu(s["Ks"]), 320, 200,
                                   calc_compensations=True)
    radii = host(got[0])
    assert_same(radii, want[0], "radii")
    vis = radii > 0
    assert vis.sum() > 100
    assert_same(host(got[1])[vis], want[1][vis], "means2d")
    assert_same(host(got[2])[vis], want[2][vis], "depths")
    assert_same(host(got[3])[vis], want[3][vis], "conics")
    assert_same(host(got[4])[vis], want[4][vis], "compensations")


@pytest.mark.parametrize("mode,C", [("covars", 1), ("covars", 3), ("quats", 2)])
def test_projection_bwd_close(device, mode, C):
    s = opaque_scene(5000, W=320, H=200, n_cams=C, seed=10 + C)
    cov = s["covars6"] if mode == "covars" else None
    q = s["quats"] if mode == "quats" else None
    sc = s["scales"] if mode == "quats" else None
    tm = cu(s["means"]).requires_grad_(True)
    tc = None if cov is None else cu(cov).requires_grad_(True)
    tq = None if q is None else cu(q).requires_grad_(True)
    ts_ = None if sc is None else cu(sc).requires_grad_(True)
    tv = cu(s["viewmats"]).requires_grad_(True)
    radii, m2, dp, cn, cp = W.fully_fused_projection(tm, tc, tq, ts_, tv, cu(s["Ks"]), 320, 200,
                                                      calc_compensations=True)
    g = np.random.Generator(np.random.PCG64(5))
    vis = host(radii) > 0
    vm2 = (g.normal(0, 1, host(m2).shape) * vis[..., None]).astype(np.float32)
    vdp = (g.normal(0, 1, host(dp).shape) * vis).astype(np.float32)
    vcn = (g.normal(0, 1, host(cn).shape) * vis[..., None]).astype(np.float32)
    vcp = (g.normal(0, 1, host(cp).shape) * vis).astype(np.float32)
    torch.autograd.backward([m2, dp, cn, cp], [cu(vm2), cu(vdp), cu(vcn), cu(vcp)])
    d = np.float64
    want = O.proj_bwd(s["means"].astype(d), None if cov is None else cov.astype(d),
                      None if q is None else q.astype(d), None if sc is None else sc.astype(d),
                      s["viewmats"].astype(d), s["Ks"].astype(d), 320, 200, 0.3, host(radii),
                      host(cn).astype(d), host(cp).astype(d), vm2.astype(d), vdp.astype(d), vcn.astype(d),
                      vcp.astype(d), viewmats_requires_grad=True)
    pairs = [(tm.grad, want[0], "v_means"), (tv.grad, want[4], "v_viewmats")]
    if cov is not None:
        pairs.append((tc.grad, want[1], "v_covars"))
    else:
        pairs += [(tq.grad, want[2], "v_quats"), (ts_.grad, want[3], "v_scales")]
    for t, w, n in pairs:
        assert rel_l2(host(t), w) < 1e-4, (n, rel_l2(host(t), w))


@pytest.mark.parametrize("C", [1, 2])
def test_projection_covars33_matches_gathered(device, C):
    # the fused [N,3,3] gather (rendering.py:54-56) against the [N,6] path
    s = opaque_scene(4000, W=320, H=200, n_cams=C, seed=30 + C)
    g = np.random.Generator(np.random.PCG64(31))
    c33 = s["covars33"] + np.tril(g.normal(0, 1, (4000, 3, 3)), -1).astype(np.float32)  # lower part is ignored
    tm6, tm33 = cu(s["means"]).requires_grad_(True), cu(s["means"]).requires_grad_(True)
    tc6, tc33 = cu(s["covars6"]).requires_grad_(True), cu(c33).requires_grad_(True)
    a = W.fully_fused_projection(tm6, tc6, None, None, cu(s["viewmats"]), cu(s["Ks"]), 320, 200,
                                 calc_compensations=True)
    b = W.fully_fused_projection_covars33(tm33, tc33, cu(s["viewmats"]), cu(s["Ks"]), 320, 200,
                                          calc_compensations=True)
    assert_same(host(a[0]), host(b[0]), "radii")
    vis = host(a[0]) > 0  # outputs of culled primitives are left unwritten, as in the reference
    for x, y, n in zip(a[1:], b[1:], ("means2d", "depths", "conics", "comp")):
        assert_same(host(x)[vis], host(y)[vis], n)
    outs = []
    for o in (a, b):
        gg = np.random.Generator(np.random.PCG64(32))
        grads = [cu((gg.normal(0, 1, host(t).shape) * (vis if t.dim() == 2 else vis[..., None])).astype(np.float32))
                 for t in o[1:]]
        torch.autograd.backward(list(o[1:]), grads)
    np.testing.assert_allclose(host(tm33.grad), host(tm6.grad), rtol=1e-6, atol=1e-6)
    G33 = host(tc33.grad)
    want = np.zeros_like(G33)
    tri = ([0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2])
    want[:, tri[0], tri[1]] = host(tc6.grad)
    np.testing.assert_allclose(G33, want, rtol=1e-6, atol=1e-6)
    assert (G33[:, [1, 2, 2], [0, 0, 1]] == 0).all()


# ------------------------------------------------------- isect / sort / offsets
@pytest.mark.parametrize("C", [1, 2])
def test_isect_and_offsets_exact(device, C):
    s = opaque_scene(30000, W=500, H=300, n_cams=C, seed=20 + C)
    r, m2, dp, cn, _ = O.proj_fwd(s["means"], s["covars6"], None, None, s["viewmats"], s["Ks"], 500, 300)
    tw, th = (500 + 15) // 16, (300 + 15) // 16
    want_tpg, want_ids, want_flat = O.isect_tiles(m2, r, dp, 16, tw, th)
    tpg, ids, flat = W.isect_tiles(cu(m2), cu(r), cu(dp), 16, tw, th)
    assert_same(host(tpg), want_tpg, "tiles_per_gauss")
    assert_same(host(ids), want_ids, "isect_ids")
    assert_same(host(flat), want_flat, "flatten_ids")
    assert want_ids.size > 10000
    off = W.isect_offset_encode(ids, C, tw, th)
    assert_same(host(off), O.isect_offset_encode(want_ids, C, tw, th), "offsets")


def test_isect_unsorted_and_edges(device):
    # primitives straddling every image border (negative tile coordinates saturate to 0)
    means2d = np.array([[[-20.0, -20.0], [5.0, 5.0], [499.0, 299.0], [250.0, -3.0], [-1.0, 150.0],
                         [510.0, 310.0]]], np.float32)
    radii = np.array([[30, 3, 40, 5, 2, 20]], np.int32)
    depths = np.array([[1.0, 2.0, 3.0, 4.0, 5.0, 0.5]], np.float32)
    tw, th = 32, 19
    for sort in (False, True):
        want = O.isect_tiles(means2d, radii, depths, 16, tw, th, sort=sort)
        got = W.isect_tiles(cu(means2d), cu(radii), cu(depths), 16, tw, th, sort=sort)
        for a, b, n in zip(got, want, ("tpg", "ids", "flat")):
            assert_same(host(a), b, f"{n} sort={sort}")


@pytest.mark.parametrize("mode", ["depth_first", "onesweep"])
def test_sort_status_reaches_the_next_host_read(device, mode):
    """A nonzero status word left by an earlier sort (what a look-back wait
    that gave up leaves; set from the host here) makes the next host read of
    the count info raise, then is cleared: the call after it gives the
    oracle's lists again.  depth_first reads the info from pinned memory,
    onesweep from the device."""
    s = opaque_scene(3000, W=200, H=120, n_cams=1, seed=7)
    r, m2, dp, _, _ = O.proj_fwd(s["means"], s["covars6"], None, None, s["viewmats"], s["Ks"], 200, 120)
    tw, th = (200 + 15) // 16, (120 + 15) // 16
    want = O.isect_tiles(m2, r, dp, 16, tw, th)
    gm2, gr, gdp = cu(m2), cu(r), cu(dp)
    before = W.isect_sort_mode()
    W.set_isect_sort_mode(mode)
    try:
        W._late_status(gm2.device).fill_(1)
        with pytest.raises(RuntimeError, match="look-back"):
            W.isect_tiles(gm2, gr, gdp, 16, tw, th)
        assert int(W._late_status(gm2.device).item()) == 0
        got = W.isect_tiles(gm2, gr, gdp, 16, tw, th)
    finally:
        W.set_isect_sort_mode(before)
    for a, b, n in zip(got, want, ("tpg", "ids", "flat")):
        assert_same(host(a), b, f"{n} {mode}")


@pytest.mark.parametrize("n", [0, 1, 63, 4095, 4096, 6143, 6144, 6145, 12289, 100003])
@pytest.mark.parametrize("end_bit", [7, 33, 46, 64])
def test_radix_sort_stable(device, n, end_bit):
    g = np.random.Generator(np.random.PCG64(n * 131 + end_bit))
    # few distinct keys -> many ties, so stability is visible
    keys = g.integers(-(1 << 62), 1 << 62, size=max(n // 7, 1), dtype=np.int64)[g.integers(0, max(n // 7, 1), n)]
    keys ^= g.integers(0, 4, n, dtype=np.int64) << 40
    vals = np.arange(n, dtype=np.int32)
    want_k, want_v = O.sort_pairs(keys, vals, 0, end_bit)
    got_k, got_v = W.radix_sort_pairs(cu(keys), cu(vals), 0, end_bit, check_status=True)
    assert_same(host(got_k), want_k, "keys")
    assert_same(host(got_v), want_v, "values")


@pytest.mark.parametrize("n", [5000, 100003])
@pytest.mark.parametrize("const_bytes", [(3,), (1, 3), (0, 1, 2, 3)])
def test_radix_sort_skips_constant_passes(device, n, const_bytes):
    """Leaving out passes over digits every key shares gives the full sort's
    result (checked against the oracle and against no skipping)."""
    g = np.random.Generator(np.random.PCG64(n + 7 * len(const_bytes)))
    keys = g.integers(0, 1 << 46, size=n, dtype=np.int64) & ~np.int64(0x3F)  # ties in the low bits
    for b in const_bytes:
        keys = (keys & ~np.int64(0xFF << (8 * b))) | np.int64(0x5A << (8 * b))
    vals = np.arange(n, dtype=np.int32)
    mask = sum(1 << b for b in const_bytes)
    want_k, want_v = O.sort_pairs(keys, vals, 0, 46)
    got_k, got_v = W.radix_sort_pairs(cu(keys), cu(vals), 0, 46, check_status=True, skip_passes=mask)
    assert_same(host(got_k), want_k, "keys")
    assert_same(host(got_v), want_v, "values")
    full_k, full_v = W.radix_sort_pairs(cu(keys), cu(vals), 0, 46)
    assert torch.equal(got_k, full_k) and torch.equal(got_v, full_v)


@pytest.mark.parametrize("end_bit,const_byte", [(46, None), (33, 3), (64, 2)])
def test_radix_sort_large_tiles(device, end_bit, const_byte):
    """Sorts of 4M keys and more run with the larger tiles; passes over a
    digit every key shares take the copy-through path on the device."""
    n = (1 << 22) + 54321
    g = np.random.Generator(np.random.PCG64(end_bit))
    keys = g.integers(-(1 << 62), 1 << 62, size=n // 5, dtype=np.int64)[g.integers(0, n // 5, n)]
    if const_byte is not None:
        keys = (keys & ~np.int64(0xFF << (8 * const_byte))) | np.int64(0x3C << (8 * const_byte))
    vals = np.arange(n, dtype=np.int32)
    want_k, want_v = O.sort_pairs(keys, vals, 0, end_bit)
    got_k, got_v = W.radix_sort_pairs(cu(keys), cu(vals), 0, end_bit, check_status=True)
    assert_same(host(got_k), want_k, "keys")
    assert_same(host(got_v), want_v, "values")


def test_isect_depth_range(device):
    g = np.random.Generator(np.random.PCG64(3))
    C, N = 2, 5000
    radii = g.integers(-1, 5, (C, N)).astype(np.int32)
    depths = g.uniform(2.5, 5.5, (C, N)).astype(np.float32)
    depths[radii <= 0] = 1e9  # culled entries must not widen the range
    means2d = g.uniform(0, 200, (C, N, 2)).astype(np.float32)
    info = torch.zeros(2, dtype=torch.int64, device="cuda")
    m_d, r_d, d_d = cu(means2d), cu(radii), cu(depths)  # held: the call takes raw pointers
    tpg = torch.empty((C, N), dtype=torch.int32, device="cuda")
    cum = torch.empty((C * N,), dtype=torch.int64, device="cuda")
    L = W.lib()
    ws_bytes = int(L.ubs_isect_count_workspace_bytes(C, N))
    ws = torch.empty((ws_bytes,), dtype=torch.uint8, device="cuda")
    L.check(L.ubs_isect_tiles_count_range(C, N, W.ptr(m_d), W.ptr(r_d), W.ptr(d_d), 16, 13, 13, W.ptr(tpg),
                                          W.ptr(cum), W.ptr(info), W._vp_at(info, 1), W.ptr(ws), ws_bytes,
                                          W.stream_of(info)), "count_range")
    n_isects, packed = (int(x) for x in info.cpu())
    assert n_isects == int(tpg.sum())
    bits = depths[radii > 0].view(np.uint32)
    assert (~packed) & 0xFFFFFFFF == int(bits.min()) and (packed >> 32) & 0xFFFFFFFF == int(bits.max())
    # 2.5 .. 5.5 share their top byte (0x40) only
    assert W.constant_depth_passes(packed) == 0b1000


# --------------------------------------------------------------------- raster
def _raster_inputs(scene_fn, W_, H_, C=1, ts=16, channels=3, **kw):
    s = scene_fn(W=W_, H=H_, n_cams=C, **kw)
    r, m2, dp, cn, _ = O.proj_fwd(s["means"], s["covars6"], None, None, s["viewmats"], s["Ks"], W_, H_)
    tw, th = (W_ + ts - 1) // ts, (H_ + ts - 1) // ts
    _, ids, flat = O.isect_tiles(m2, r, dp, ts, tw, th)
    off = O.isect_offset_encode(ids, C, tw, th)
    op = np.ascontiguousarray(np.broadcast_to(s["opac"], (C, s["opac"].size)))
    be = np.ascontiguousarray(np.broadcast_to(s["beta"], (C, s["beta"].size)))
    rgb = s["rgb"] if channels == 3 else \
        np.random.Generator(np.random.PCG64(channels)).uniform(0, 1, (s["rgb"].shape[0], channels)).astype(np.float32)
    col = np.ascontiguousarray(np.broadcast_to(rgb, (C,) + rgb.shape))
    return dict(m2=m2, cn=cn, col=col, op=op, be=be, off=off, flat=flat, tw=tw, th=th, C=C, ts=ts)


@pytest.mark.parametrize("bg", [False, True])
def test_raster_fwd_close(device, bg):
    W_, H_ = 200, 136
    x = _raster_inputs(lambda **k: opaque_scene(3000, **k), W_, H_, C=2)
    bgs = np.random.Generator(np.random.PCG64(3)).uniform(0, 1, (2, 3)).astype(np.float32) if bg else None
    want_c, want_a, want_last, mT, mS = O.raster_fwd(x["m2"], x["cn"], x["col"], x["op"], x["be"], W_, H_, 16,
                                                    x["off"], x["flat"], bgs, margins=True)
    rc, ra, last, Tf = W.raster_fwd_raw(cu(x["m2"]), cu(x["cn"]), cu(x["col"]), cu(x["op"]), cu(x["be"]),
                                        None if bgs is None else cu(bgs), None, W_, H_, 16, cu(x["off"]),
                                        cu(x["flat"]))
    safe = (mT > 1e-3) & (mS > 1e-6)  # pixels far from the T cutoff / sigma=0 branch points
    assert safe.mean() > 0.95
    np.testing.assert_allclose(host(rc)[safe], want_c[safe], atol=2e-5)
    np.testing.assert_allclose(host(ra)[..., 0][safe], want_a[..., 0][safe], atol=2e-5)
    assert_same(host(last)[safe], want_last[safe], "last_ids")
    np.testing.assert_allclose(host(Tf)[safe], 1.0 - want_a[..., 0][safe], atol=2e-5)
    assert (want_a > 0.9).mean() > 0.05  # the cutoff is exercised
    check_branch_pixels(host(rc), host(last), want_c, want_last, np.ones(safe.shape, bool), safe,
                        what=f"blend fwd bg={bg}")


def test_raster_bwd_close(device):
    W_, H_ = 160, 120
    x = _raster_inputs(lambda **k: opaque_scene(1500, opacity=(0.2, 0.9), **k), W_, H_, C=1)
    d = np.float64
    want_c, want_a, want_last, mT, mS = O.raster_fwd(*(x[k].astype(d) for k in ("m2", "cn", "col", "op", "be")),
                                                    W_, H_, 16, x["off"], x["flat"], margins=True)
    rc, ra, last, Tf = W.raster_fwd_raw(cu(x["m2"]), cu(x["cn"]), cu(x["col"]), cu(x["op"]), cu(x["be"]), None,
                                        None, W_, H_, 16, cu(x["off"]), cu(x["flat"]))
    assert (mT > 1e-3).all() and (mS > 1e-6).all(), "pick a seed without branch-point pixels"
    g = np.random.Generator(np.random.PCG64(9))
    vrc = g.normal(0, 1, host(rc).shape).astype(np.float32)
    vra = g.normal(0, 1, host(ra).shape).astype(np.float32)
    got = W.raster_bwd_raw(cu(x["m2"]), cu(x["cn"]), cu(x["col"]), cu(x["op"]), cu(x["be"]), None, None, W_, H_,
                           16, cu(x["off"]), cu(x["flat"]), ra, last, Tf, cu(vrc), cu(vra))
    want = O.raster_bwd(*(x[k].astype(d) for k in ("m2", "cn", "col", "op", "be")), W_, H_, 16, x["off"],
                        x["flat"], want_a, want_last, vrc.astype(d), vra.astype(d), T_final=1.0 - want_a[..., 0])
    for a, b, n in zip(got, want, ("v_means2d", "v_conics", "v_colors", "v_opac", "v_betas")):
        assert rel_l2(host(a), b) < 2e-3, (n, rel_l2(host(a), b))


@pytest.mark.parametrize("C,channels,ts,bg,mask", [
    (1, 1, 16, False, False), (2, 3, 8, True, False), (1, 4, 32, False, False), (1, 7, 16, True, True),
    (2, 32, 16, False, False), (1, 35, 8, False, True), (1, 3, 32, True, False), (1, 2, 16, True, False),
    (2, 10, 16, False, True), (1, 18, 8, True, False)])
def test_raster_bwd_variants(device, C, channels, ts, bg, mask):
    """Blend backward over channel counts (chunked register tiles), tile
    sizes, cameras, backgrounds and tile masks.  Upstream gradients are
    zeroed on pixels near a forward branch point (T cutoff, sigma = 0), so
    every remaining term is smooth and the FP64 oracle is the exact target."""
    W_, H_ = 136, 104
    x = _raster_inputs(lambda **k: opaque_scene(1200, opacity=(0.2, 0.95), **k), W_, H_, C=C, ts=ts,
                       channels=channels)
    g = np.random.Generator(np.random.PCG64(100 + channels))
    bgs = g.uniform(0, 1, (C, channels)).astype(np.float32) if bg else None
    masks = (g.uniform(0, 1, (C, x["th"], x["tw"])) > 0.25) if mask else None
    d = np.float64
    want_c, want_a, want_last, mT, mS = O.raster_fwd(*(x[k].astype(d) for k in ("m2", "cn", "col", "op", "be")),
                                                    W_, H_, ts, x["off"], x["flat"], backgrounds=bgs, masks=masks,
                                                    margins=True)
    tmask = None if masks is None else cu(masks.astype(np.bool_))
    rc, ra, last, Tf = W.raster_fwd_raw(cu(x["m2"]), cu(x["cn"]), cu(x["col"]), cu(x["op"]), cu(x["be"]),
                                        None if bgs is None else cu(bgs), tmask, W_, H_, ts, cu(x["off"]),
                                        cu(x["flat"]))
    safe = (mT > 1e-3) & (mS > 1e-6)
    if masks is not None:  # masked tiles blend nothing: no branch points, gradients must stay zero
        off_px = ~np.repeat(np.repeat(masks, ts, axis=1), ts, axis=2)[:, :H_, :W_]
        assert off_px.any() and safe[~off_px].mean() > 0.9
        safe |= off_px
    assert safe.mean() > 0.9
    np.testing.assert_allclose(host(rc)[safe], want_c[safe], atol=3e-5)
    # the pixels near a branch point against the float32 blend of the same lists
    c32, _, last32 = O.raster_fwd(x["m2"], x["cn"], x["col"], x["op"], x["be"], W_, H_, ts, x["off"], x["flat"],
                                  backgrounds=bgs, masks=masks)
    check_branch_pixels(host(rc), host(last), c32, last32, np.ones(safe.shape, bool), safe, atol=3e-5,
                        what=f"blend fwd C={C} channels={channels} ts={ts}")
    vrc = (g.normal(0, 1, host(rc).shape) * safe[..., None]).astype(np.float32)
    vra = (g.normal(0, 1, host(ra).shape) * safe[..., None]).astype(np.float32)
    got = W.raster_bwd_raw(cu(x["m2"]), cu(x["cn"]), cu(x["col"]), cu(x["op"]), cu(x["be"]),
                           None if bgs is None else cu(bgs), tmask, W_, H_, ts, cu(x["off"]), cu(x["flat"]), ra,
                           last, Tf, cu(vrc), cu(vra))
    want = O.raster_bwd(*(x[k].astype(d) for k in ("m2", "cn", "col", "op", "be")), W_, H_, ts, x["off"],
                        x["flat"], want_a, want_last, vrc.astype(d), vra.astype(d), backgrounds=bgs, masks=masks,
                        T_final=1.0 - want_a[..., 0])
    for a, b, n in zip(got, want, ("v_means2d", "v_conics", "v_colors", "v_opac", "v_betas")):
        assert np.abs(b).max() > 0, n
        assert rel_l2(host(a), b) < 2e-3, (n, rel_l2(host(a), b))


def test_wrong_dtypes_raise(device):
    """fp64 / int64 buffers are rejected before any launch (the reference's
    data_ptr<float>() / data_ptr<int32_t>() checks), never read as fp32."""
    C, N = 1, 16
    f = lambda *s: torch.zeros(*s, device="cuda")
    off = torch.zeros(C, 2, 2, dtype=torch.int32, device="cuda")
    flat = torch.zeros(0, dtype=torch.int32, device="cuda")
    with pytest.raises(RuntimeError, match="expected scalar type"):
        W.rasterize_to_pixels(f(C, N, 2), f(C, N, 3), f(C, N, 3), f(C, N), f(C, N), 32, 32, 16, off, flat,
                              backgrounds=f(C, 3).double())
    with pytest.raises(RuntimeError, match="expected scalar type"):
        W.rasterize_to_pixels(f(C, N, 2), f(C, N, 3), f(C, N, 3), f(C, N), f(C, N), 32, 32, 16, off.long(), flat)
    with pytest.raises(RuntimeError, match="expected scalar type"):
        W.fully_fused_projection(f(N, 3).double(), f(N, 6).double(), None, None, f(1, 4, 4).double(),
                                 f(1, 3, 3).double(), 32, 32)
    with pytest.raises(RuntimeError, match="expected scalar type"):
        W.isect_offset_encode(torch.zeros(4, dtype=torch.int32, device="cuda"), 1, 2, 2)


# ------------------------------------------------------------- activations
@pytest.mark.parametrize("D,quats", [(6, True), (7, True), (6, False)])
def test_beta_activations_vs_torch(device, D, quats):
    """The fused activation/query/rotation kernel against the torch ops it
    replaces (BetaModel's getters and render query), forward and backward,
    both in fp32 on the GPU."""
    import torch.nn.functional as F

    from gsplat.beta import quat_to_rotmat

    g = torch.Generator(device="cpu").manual_seed(D)
    N = 5000
    mk = lambda *s, sc=1.0: (torch.randn(*s, generator=g) * sc).cuda().requires_grad_(True)
    xyz, rest, sraw = mk(N, 3), mk(N, D - 3), mk(N, D, sc=8.0)  # scale 8: exercises softplus' x > 20 branch
    oraw, braw = mk(N, 1), mk(N, D - 2, sc=0.5)
    q = mk(N, 4) if quats else None
    cam = torch.tensor([0.3, -4.0, 1.0], device="cuda")
    t = 0.37
    mean, scale, opac, beta, query, rot = W.beta_activations(xyz, rest, sraw, oraw, braw, q, cam, t)
    leaves = [xyz, rest, sraw, oraw, braw] + ([q] if quats else [])
    ref_leaves = [l.detach().clone().requires_grad_(True) for l in leaves]
    rx, rr, rs, ro, rb = ref_leaves[:5]
    r_mean = torch.cat([rx, rr], -1)
    r_scale = F.softplus(rs)
    r_opac = torch.sigmoid(ro)
    r_beta = 4.0 * torch.exp(rb)
    with torch.no_grad():
        vd = rx - cam[None]
        r_query = vd / vd.norm(dim=-1, keepdim=True)
        if D == 7:
            r_query = torch.cat([r_query, torch.full((N, 1), t, device="cuda")], -1)
    assert torch.equal(mean, r_mean)
    for a, b, n in ((scale, r_scale, "scale"), (opac, r_opac, "opacity"), (beta, r_beta, "beta"),
                    (query, r_query, "query")):
        torch.testing.assert_close(a, b, rtol=2e-6, atol=1e-7, msg=n)
    outs, refs = [mean, scale, opac, beta], [r_mean, r_scale, r_opac, r_beta]
    if quats:
        r_rot = quat_to_rotmat(ref_leaves[5])
        torch.testing.assert_close(rot, r_rot, rtol=1e-5, atol=1e-6)  # O(1) entries: a few ulp
        outs.append(rot)
        refs.append(r_rot)
    else:
        assert rot is None
    grads = [torch.randn(o.shape, generator=g).cuda() for o in outs]
    torch.autograd.backward(outs, grads)
    torch.autograd.backward(refs, grads)
    for a, b in zip(leaves, ref_leaves):
        torch.testing.assert_close(a.grad, b.grad, rtol=1e-5, atol=1e-5)
    assert not query.requires_grad


@pytest.mark.parametrize("C,channels,ts", [(2, 3, 16), (1, 35, 8), (1, 4, 32)])
def test_blend_packed_records_match_unpacked(device, C, channels, ts):
    """The packed-record gathers (default) against the four-array gathers."""
    W_, H_ = 136, 104
    x = _raster_inputs(lambda **k: opaque_scene(1200, opacity=(0.2, 0.95), **k), W_, H_, C=C, ts=ts,
                       channels=channels)
    ins = [cu(x[k]) for k in ("m2", "cn", "col", "op", "be")]
    g = np.random.Generator(np.random.PCG64(7))
    bgs = cu(g.uniform(0, 1, (C, channels)).astype(np.float32))
    res = {}
    try:
        for packed in (False, True):
            W.set_blend_packed_records(packed)
            fwd = W.raster_fwd_raw(*ins, bgs, None, W_, H_, ts, cu(x["off"]), cu(x["flat"]))
            vrc = torch.from_numpy(np.random.Generator(np.random.PCG64(8)).normal(0, 1, fwd[0].shape).astype(
                np.float32)).cuda()
            vra = torch.ones_like(fwd[1])
            bwd = W.raster_bwd_raw(*ins, bgs, None, W_, H_, ts, cu(x["off"]), cu(x["flat"]), fwd[1], fwd[2], fwd[3],
                                   vrc, vra)
            res[packed] = (fwd, bwd)
    finally:
        W.set_blend_packed_records(True)
    for a, b in zip(res[False][0], res[True][0]):
        assert torch.equal(a, b)
    for a, b, n in zip(res[False][1], res[True][1], ("v_means2d", "v_conics", "v_colors", "v_opac", "v_betas")):
        assert rel_l2(host(b), host(a).astype(np.float64)) < 1e-5, n


@pytest.mark.parametrize("C,channels,ts,mask", [(2, 3, 16, False), (1, 3, 8, True), (3, 4, 32, False),
                                                (1, 7, 16, False)])
def test_blend_tile_order_same_results(device, C, channels, ts, mask):
    """Tile blocks dealt heaviest first (the default, the workspace of
    ubs_rasterize_workspace_bytes_tiles) against raster order: the forward
    bit for bit, the backward within the rounding of its atomics."""
    W_, H_ = 232, 152
    x = _raster_inputs(lambda **k: opaque_scene(2500, opacity=(0.2, 0.95), **k), W_, H_, C=C, ts=ts,
                       channels=channels)
    ins = [cu(x[k]) for k in ("m2", "cn", "col", "op", "be")]
    g = np.random.Generator(np.random.PCG64(11))
    masks = cu(g.uniform(0, 1, (C, x["th"], x["tw"])) < 0.8) if mask else None
    bgs = cu(g.uniform(0, 1, (C, channels)).astype(np.float32))
    res = {}
    try:
        for ordered in (False, True):
            W.set_blend_tile_order(ordered)
            fwd = W.raster_fwd_raw(*ins, bgs, masks, W_, H_, ts, cu(x["off"]), cu(x["flat"]))
            vrc = cu(np.random.Generator(np.random.PCG64(8)).normal(0, 1, tuple(fwd[0].shape)).astype(np.float32))
            vra = torch.ones_like(fwd[1])
            bwd = W.raster_bwd_raw(*ins, bgs, masks, W_, H_, ts, cu(x["off"]), cu(x["flat"]), fwd[1], fwd[2],
                                   fwd[3], vrc, vra)
            res[ordered] = (fwd, bwd)
    finally:
        W.set_blend_tile_order(True)
    for a, b in zip(res[False][0], res[True][0]):
        assert torch.equal(a, b)
    for a, b, n in zip(res[False][1], res[True][1], ("v_means2d", "v_conics", "v_colors", "v_opac", "v_betas")):
        assert rel_l2(host(b), host(a).astype(np.float64)) < 1e-6, n


@pytest.mark.parametrize("C,W_,H_,ts", [(1, 232, 152, 16), (3, 200, 136, 8), (2, 1000, 600, 32)])
def test_tile_order_is_per_xcd_heaviest_first(device, C, W_, H_, ts):
    """The order the forward writes into its workspace, read back: for each
    of the 8 XCDs exactly the tiles its 2x2 blocks hold (the XCD the raster
    order gives them), heaviest first to the 256-bucket width of the order
    kernel, then padding (~0)."""
    x = _raster_inputs(lambda **k: opaque_scene(3000, **k), W_, H_, C=C, ts=ts)
    ins = [cu(x[k]) for k in ("m2", "cn", "col", "op", "be")]
    out = W._raster_fwd(*ins, None, None, W_, H_, ts, cu(x["off"]), cu(x["flat"]), False)
    ws = out[5]
    tw, th = x["tw"], x["th"]
    L = W.lib()
    base = int(L.ubs_rasterize_workspace_bytes(C, x["op"].shape[1], 0))
    nbx, nby = (tw + 1) // 2, (th + 1) // 2
    slots = (C * nbx * nby + 7) // 8 * 4
    order = host(ws[base:base + 4 * 8 * slots].view(torch.int32)).view(np.uint32).reshape(8, slots)
    tiles = C * tw * th
    off = x["off"].reshape(-1).astype(np.int64)
    cnt = np.diff(np.append(off, x["flat"].size))
    work = cnt + 64
    t = np.arange(tiles)
    cam, rem = t // (tw * th), t % (tw * th)
    xcd = (cam * nbx * nby + (rem // tw) // 2 * nbx + (rem % tw) // 2) % 8
    width = (work.max() + 1) / 256.0
    seen = np.zeros(tiles, np.int64)
    for k in range(8):
        lst = order[k]
        real = lst[lst != 0xFFFFFFFF]
        assert (lst[real.size:] == 0xFFFFFFFF).all(), k  # padding only after the tiles
        assert (xcd[real] == k).all(), k
        seen[real] += 1
        assert set(real.tolist()) == set(np.nonzero(xcd == k)[0].tolist()), k
        b = np.floor(work[real] * 256 / (work.max() + 1))
        assert (np.diff(b) <= 0).all(), k  # heaviest bucket first
    assert (seen == 1).all()
    assert width > 0


@pytest.mark.parametrize("C,channels,ts,ckpt", [(2, 3, 16, False), (1, 3, 16, True), (1, 7, 8, False),
                                                (1, 40, 16, True)])
def test_blend_backward_reuses_forward_workspace(device, C, channels, ts, ckpt):
    """ubs_rasterize_to_pixels_bwd_reuse (the forward's packed records and
    tile order, a gradient-only workspace) against the backward that packs
    and orders again: same gradients within the rounding of its atomics, with
    and without transmittance checkpoints, narrow and wide records."""
    W_, H_ = 200, 136
    x = _raster_inputs(lambda **k: opaque_scene(2000, opacity=(0.2, 0.95), **k), W_, H_, C=C, ts=ts,
                       channels=channels)
    ins = [cu(x[k]) for k in ("m2", "cn", "col", "op", "be")]
    bgs = cu(np.random.Generator(np.random.PCG64(12)).uniform(0, 1, (C, channels)).astype(np.float32))
    out = W._raster_fwd(*ins, bgs, None, W_, H_, ts, cu(x["off"]), cu(x["flat"]), ckpt)
    rc, ra, last, Tf, T_ckpt, ws = out
    assert ws is not None and (T_ckpt is not None) == ckpt
    vrc = cu(np.random.Generator(np.random.PCG64(13)).normal(0, 1, tuple(rc.shape)).astype(np.float32))
    vra = cu(np.random.Generator(np.random.PCG64(14)).normal(0, 1, tuple(ra.shape)).astype(np.float32))
    args = (*ins, bgs, None, W_, H_, ts, cu(x["off"]), cu(x["flat"]), ra, last, Tf, vrc, vra)
    plain = W.raster_bwd_raw(*args, T_ckpt=T_ckpt)
    reuse = W.raster_bwd_raw(*args, T_ckpt=T_ckpt, fwd_workspace=ws)
    for a, b, n in zip(plain, reuse, ("v_means2d", "v_conics", "v_colors", "v_opac", "v_betas")):
        assert rel_l2(host(b), host(a).astype(np.float64)) < 1e-6, n


@pytest.mark.parametrize("case", ["random", "dense_ties", "dense_4_bytes", "two_cams"])
def test_segmented_depth_sort_matches_full_sort(device, case):
    """Every sort mode -- depth-first emission + onesweep on the camera/tile
    bits, onesweep on the camera/tile bits + per-tile depth sort, onesweep on
    every byte -- gives the oracle's ids and values bit for bit: segments
    that fit LDS and segments longer than 8192 (HBM chunks, odd and even
    pass counts), equal depths (stability), skipped constant depth bytes,
    culled primitives, two cameras."""
    g = np.random.Generator(np.random.PCG64({"random": 1, "dense_ties": 2, "dense_4_bytes": 3, "two_cams": 4}[case]))
    C = 2 if case == "two_cams" else 1
    W_, H_, ts = 160, 128, 16
    tw, th = W_ // ts, H_ // ts
    N = 30_000
    if case.startswith("dense"):  # most primitives cover the centre tiles: segments of ~10k entries
        m2 = g.normal((80.0, 64.0), 6.0, (C, N, 2))
        radii = g.integers(8, 24, (C, N))
    else:
        m2 = g.uniform(0, (W_, H_), (C, N, 2))
        radii = g.integers(0, 20, (C, N))
    if case == "dense_ties":  # 40 distinct depths in [2, 4): long runs of equal keys, 3 depth passes
        depths = g.choice(np.linspace(2.0, 3.9, 40), (C, N))
    elif case == "dense_4_bytes":  # depths spanning exponents: no constant depth byte
        depths = np.exp(g.uniform(-8, 8, (C, N)))
    else:
        depths = g.uniform(2.0, 6.0, (C, N))
    m2, radii, depths = m2.astype(np.float32), radii.astype(np.int32), depths.astype(np.float32)
    _, want_ids, want_flat = O.isect_tiles(m2, radii, depths, ts, tw, th)
    if case.startswith("dense"):
        off = O.isect_offset_encode(want_ids, C, tw, th).reshape(-1)
        assert np.diff(np.append(off, want_ids.size)).max() > 8192  # the chunked HBM path
    got = {}
    before = W.isect_sort_mode()
    for mode in ("depth_first", "segmented", "onesweep"):
        W.set_isect_sort_mode(mode)
        try:
            _, ids, flat = W.isect_tiles(cu(m2), cu(radii), cu(depths), ts, tw, th)
            _, _, _, offs = W.isect_tiles_offsets(cu(m2), cu(radii), cu(depths), ts, tw, th)
        finally:
            W.set_isect_sort_mode(before)
        assert_same(host(ids), want_ids, f"ids {mode}")
        assert_same(host(flat), want_flat, f"flat {mode}")
        if offs is not None:  # offsets handed out on the way are those of the final ids
            assert_same(host(offs), O.isect_offset_encode(want_ids, C, tw, th), f"offsets {mode}")


# ------------------------------------------------------------- camera centres
def test_camera_centers_match_the_fused_projection(device):
    """ubs_camera_centers gives -(R^T t) (float64 reference, float32
    tolerance), and the fused projection given no cam_pos takes the same
    centres: every output equals the one with cam_pos = camera_centers(...)
    bit for bit, for 3 views of a 6-D scene."""
    from gsplat import beta

    vms, Ks, _ = scenes.make_cameras(3, 200, 120, seed=4, radius=(3.0, 5.0))
    want = np.stack([-(v[:3, :3].astype(np.float64).T @ v[:3, 3].astype(np.float64)) for v in vms])
    got = host(W.camera_centers(cu(vms)))
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
    params = beta.BetaParams.from_scene(scenes.make_scene(3000, 6, seed=5), torch.device(device))
    args = (params.xyz, params.mean_rest, params.scale_raw, params.l_triangle, params.opacity_raw, params.beta_raw,
            params.quats, cu(vms), cu(Ks), 200, 120)
    with torch.no_grad():
        a = W.beta_project(*args)
        b = W.beta_project(*args, cam_pos=W.camera_centers(cu(vms)))
    for x, y, n in zip(a, b, ("radii", "means2d", "depths", "conics", "opacities", "betas")):
        vis = host(a[0]) > 0
        xa, ya = host(x), host(y)
        if n != "radii":  # (culled entries' 2-D outputs are left untouched)
            xa, ya = xa[vis], ya[vis]
        assert_same(xa, ya, n)
