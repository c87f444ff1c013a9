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
