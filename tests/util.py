"""Shared helpers of the test suite: synthetic inputs (numpy) and comparisons."""
from __future__ import annotations

import numpy as np

from gsplat import scenes
from oracle import ubs_oracle as O


def softplus(x):
    return np.logaddexp(0.0, x)


def sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def activations(sc: scenes.Scene, dt=np.float32):
    """Activated leaf parameters (FP64 math, cast to dt)."""
    x = {k: getattr(sc, k).astype(np.float64) for k in
         ("xyz", "mean_rest", "scale_raw", "rot", "l_triangle", "opacity_raw", "beta_raw", "rgb")}
    return dict(
        scale=softplus(x["scale_raw"]).astype(dt),
        opac=sigmoid(x["opacity_raw"]).astype(dt),
        beta=(4.0 * np.exp(x["beta_raw"])).astype(dt),
        mean=np.concatenate([x["xyz"], x["mean_rest"]], 1).astype(dt),
        rot=x["rot"].astype(dt),
        ltri=x["l_triangle"].astype(dt),
        rgb=x["rgb"].astype(dt),
        xyz=x["xyz"].astype(dt),
    )


def query_for(xyz: np.ndarray, viewmat: np.ndarray, D: int, t: float = 0.0, dt=np.float32):
    cc = scenes.camera_center(viewmat.astype(np.float64))
    vd = xyz.astype(np.float64) - cc
    vd /= np.linalg.norm(vd, axis=1, keepdims=True)
    if D == 7:
        vd = np.concatenate([vd, np.full((vd.shape[0], 1), t)], 1)
    return vd.astype(dt)


def conditioned_scene(N, D=6, W=256, H=256, seed=0, n_cams=1, dt=np.float32):
    """Scene -> oracle covariance + conditioning (float32), cameras.

    Returns dict(means [N,3], covars [N,3,3], opac [N], beta [N], rgb [N,3],
    viewmats [C,4,4], Ks [C,3,3], ts [C], act=activations, cov=DxD covariances)."""
    sc = scenes.make_scene(N, D, seed=seed)
    vms, Ks, ts = scenes.make_cameras(n_cams, W, H, seed=seed)
    a = activations(sc, dt)
    ri, rj = scenes.rest_pattern(D)
    cov = O.covar_fwd(a["rot"], a["scale"], a["ltri"], ri, rj)
    q = query_for(a["xyz"], vms[0], D, float(ts[0]), dt)
    m, V, o = O.cond_fwd(a["mean"], cov, a["opac"], a["beta"][:, 1:], q)
    return dict(means=m, covars=V, opac=np.ascontiguousarray(o[:, 0]), beta=np.ascontiguousarray(a["beta"][:, 0]),
                rgb=a["rgb"], viewmats=vms.astype(dt), Ks=Ks.astype(dt), ts=ts, act=a, cov=cov, query=q,
                rest=(ri, rj), scene=sc)


def opaque_scene(N, W=128, H=128, seed=0, n_cams=1, opacity=(0.6, 0.99), beta=(0.5, 6.0), scale=(0.02, 0.12)):
    """3-D primitives with high opacity (exercises the T <= 1e-4 cutoff and the
    0.999 clamp): means [N,3], covars6 [N,6], quats, scales, opac, beta, rgb."""
    g = np.random.Generator(np.random.PCG64(seed + 77))
    means = g.uniform(-1.0, 1.0, (N, 3))
    q = g.normal(0.0, 1.0, (N, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    s = g.uniform(scale[0], scale[1], (N, 3))
    R = scenes.quat_to_rotmat_np(q)
    M = R * s[:, None, :]
    cov = M @ np.transpose(M, (0, 2, 1))
    tri = ([0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2])
    cov6 = cov[:, tri[0], tri[1]]
    vms, Ks, ts = scenes.make_cameras(n_cams, W, H, seed=seed)
    return dict(means=means.astype(np.float32), covars6=np.ascontiguousarray(cov6.astype(np.float32)),
                covars33=cov.astype(np.float32), quats=q.astype(np.float32), scales=s.astype(np.float32),
                opac=g.uniform(opacity[0], opacity[1], N).astype(np.float32),
                beta=g.uniform(beta[0], beta[1], N).astype(np.float32),
                rgb=g.uniform(0, 1, (N, 3)).astype(np.float32), viewmats=vms.astype(np.float32),
                Ks=Ks.astype(np.float32))


def assert_same(a, b, what=""):
    """Bit-exact equality (NaNs at the same places count as equal)."""
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    if a.dtype.kind == "f":
        both_nan = np.isnan(a) & np.isnan(b)
        ok = (a == b) | both_nan
    else:
        ok = a == b
    if not ok.all():
        idx = np.argwhere(~ok)[:5]
        raise AssertionError(f"{what}: {int((~ok).sum())} of {a.size} differ; first at {idx.tolist()}: "
                             f"{[a[tuple(i)] for i in idx]} vs {[b[tuple(i)] for i in idx]}")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def near_singular_2d_cases():
    """Projection inputs whose blurred 2-D covariance is [[c, c01], [c01, c]]
    exactly, from near-singular to singular and indefinite.

    The camera is the identity with fx = fy = z = 4 and the means sit on the
    optical axis, so the perspective Jacobian is the identity and the 2-D
    covariance equals the 3-D covariance's upper-left block bit for bit; the
    blur adds 0.3 in fp32.  With c01 = c - k ulp(c) the exact determinant is
    k ulp(c) (2c - k ulp(c)): positive and tiny for k >= 1 (far below what
    fp32 resolves in c^2), zero for k = 0, negative for k < 0.  Returns
    (means [M,3], cov6 [M,6], viewmats [1,4,4], Ks [1,3,3], W, H, c [M],
    c01 [M]) in float32."""
    W = H = 64
    K = np.array([[[4.0, 0.0, 32.0], [0.0, 4.0, 32.0], [0.0, 0.0, 1.0]]], np.float32)
    vm = np.eye(4, dtype=np.float32)[None]
    blur = np.float32(0.3)
    rows, cs, c01s = [], [], []
    for lam in (16.0, 64.0, 256.0, 1024.0):
        s = np.float32(lam / 2)
        c = np.float32(s + blur)
        ulp = np.float32(np.spacing(c))
        for sign in (1.0, -1.0):
            for k in (-3, -1, 0, 1, 2, 5, 40, 1000, 30000, 200000):
                c01 = np.float32(sign * np.float64(c - np.float32(k) * ulp))
                if abs(np.float64(c01)) >= 2 ** 24 * np.float64(ulp):
                    continue
                # the 3-D entry whose blurred 2-D counterpart is c01 (no blur off-diagonal)
                rows.append([s, c01, 0.0, s, 0.0, s])
                cs.append(c)
                c01s.append(c01)
    M = len(rows)
    means = np.tile(np.array([[0.0, 0.0, 4.0]], np.float32), (M, 1))
    return (means, np.asarray(rows, np.float32), vm, K, W, H, np.asarray(cs, np.float32),
            np.asarray(c01s, np.float32))


def expected_conic_2d(c, c01):
    """Per case of near_singular_2d_cases: (exact determinant as a Fraction,
    the exact inverse's three entries rounded to float32, the fp32 formula's
    determinant and inverse, whether the FP64 branch applies)."""
    from fractions import Fraction

    c, c01 = np.float32(c), np.float32(c01)
    det_exact = Fraction(float(c)) * Fraction(float(c)) - Fraction(float(c01)) * Fraction(float(c01))
    det32 = np.float32(np.float32(c * c) - np.float32(c01 * c01))
    branch = bool(det32 < np.float32(np.float32(0.00390625) * np.float32(c * c)))
    exact = None
    if det_exact > 0:
        exact = np.array([float(Fraction(float(c)) / det_exact), float(-Fraction(float(c01)) / det_exact),
                          float(Fraction(float(c)) / det_exact)]).astype(np.float32)
    fp32 = None
    if det32 > 0:
        invd = np.float32(np.float32(1.0) / det32)
        fp32 = np.array([c * invd, -c01 * invd, c * invd], np.float32)
    return det_exact, exact, det32, fp32, branch


def cross_correlated_6d(N, seed=0):
    """6-D primitives with strong spatial/view cross-covariance and unequal
    Beta factors below 4 (beta_adj = beta / 4 < 1 per query dim), for which
    the conditioned covariance's mirrored upper triangle is often indefinite
    (about one in five): means [N,6], covars [N,6,6], opacities [N,1],
    betas [N,3], query [N,3], float32."""
    g = np.random.Generator(np.random.PCG64(seed))
    L = np.tril(g.normal(0.0, 1.0, (N, 6, 6)))
    s = np.exp(g.normal(-2.5, 0.7, (N, 6)))
    L = L * s[:, :, None]
    idx = np.arange(6)
    L[:, idx, idx] = np.abs(L[:, idx, idx]) + 1e-3 * s
    cov = L @ np.transpose(L, (0, 2, 1))
    means = np.concatenate([g.uniform(-0.3, 0.3, (N, 3)), g.uniform(-1.0, 1.0, (N, 3))], 1)
    opac = g.uniform(0.3, 0.9, (N, 1))
    betas = g.uniform(0.2, 3.9, (N, 3))
    q = g.normal(0.0, 1.0, (N, 3))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    f = lambda a: np.ascontiguousarray(a, np.float32)
    return f(means), f(cov), f(opac), f(betas), f(q)


def blurred_det_ratio(means3, V33, viewmat, K, W, H, eps2d=0.3):
    """det / (c00 c11) of the blurred 2-D covariance of each primitive (mean
    means3 [N,3], covariance the mirrored upper triangle of V33 [N,3,3]) in
    one camera, evaluated in FP64 with the projection's Jacobian and clamp;
    the projection's FP64 branch applies below 2^-8."""
    V = np.asarray(V33, np.float64)
    S = np.triu(V) + np.transpose(np.triu(V, 1), (0, 2, 1))
    vm, K = np.asarray(viewmat, np.float64), np.asarray(K, np.float64)
    R, t = vm[:3, :3], vm[:3, 3]
    pc = np.asarray(means3, np.float64) @ R.T + t
    Sc = R @ S @ R.T
    fx, fy, cx, cy = K[0, 0], K[1, 1], K[0, 2], K[1, 2]
    x, y, z = pc[:, 0], pc[:, 1], pc[:, 2]
    tfx, tfy = 0.5 * W / fx, 0.5 * H / fy
    tx = z * np.minimum((W - cx) / fx + 0.3 * tfx, np.maximum(-(cx / fx + 0.3 * tfx), x / z))
    ty = z * np.minimum((H - cy) / fy + 0.3 * tfy, np.maximum(-(cy / fy + 0.3 * tfy), y / z))
    J = np.zeros((pc.shape[0], 2, 3))
    J[:, 0, 0], J[:, 0, 2] = fx / z, -fx * tx / z ** 2
    J[:, 1, 1], J[:, 1, 2] = fy / z, -fy * ty / z ** 2
    c2 = J @ Sc @ np.transpose(J, (0, 2, 1))
    c00, c11, c01 = c2[:, 0, 0] + eps2d, c2[:, 1, 1] + eps2d, c2[:, 0, 1]
    return (c00 * c11 - c01 * c01) / (c00 * c11)


# Largest share of the checked pixels whose blend may take the other side of
# a branch point (the T <= 1e-4 stop, the sigma < 0 skip) than the float32
# oracle does: the GPU's (1 - sigma)^beta is exp2(beta log2(1 - sigma)) on the
# hardware transcendentals, the oracle's is libm pow, and the two differ in
# the last bits of alpha and T.
BRANCH_FLIP_FRAC = 5e-4


def check_branch_pixels(got_c, got_last, want_c, want_last, checked, safe, atol=2e-5,
                        max_frac=BRANCH_FLIP_FRAC, what="blend"):
    """Bound the pixels that the close comparisons leave out.

    got_* / want_* are the GPU blend and the float32 oracle's blend of the
    same lists (colour [..., H, W, D], last ids [..., H, W]); `checked` is the
    pixel mask of the comparison, `safe` the pixels away from a branch point.
    Every checked pixel near a branch point must agree to `atol` in colour and
    exactly in last id, except at most `max_frac` of the checked pixels (and
    at least one): a decision taken the other way.  Returns the counts."""
    checked = np.asarray(checked, bool)
    near = checked & ~np.asarray(safe, bool)
    dc = np.abs(np.asarray(got_c, np.float64) - np.asarray(want_c, np.float64)).max(axis=-1)
    bad = near & ((dc > atol) | (np.asarray(got_last) != np.asarray(want_last)))
    n_checked, n_near, n_bad = int(checked.sum()), int(near.sum()), int(bad.sum())
    allowed = max(1, int(max_frac * n_checked))
    print(f"{what}: {n_near} of {n_checked} checked pixels near a branch point "
          f"({n_near / max(n_checked, 1):.2%}); {n_bad} take the other branch (allowed {allowed})", flush=True)
    assert n_bad <= allowed, (what, n_bad, allowed)
    return n_checked, n_near, n_bad
