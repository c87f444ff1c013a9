"""Synthetic Beta-splat scenes and cameras for the benchmark workloads.

Draw order (seeded ``numpy.random.Generator(PCG64(seed))``, all draws in
FP64, cast to FP32 at the end):

1. xyz ~ U[-1, 1]^3
2. spatial log-scales ~ N(-3.5, 0.5)            -> scale = exp(ls)
3. quaternion ~ normalise(N(0, I_4))            -> rotation R(q)
4. angular (view-direction) log-scales ~ N(0, 0.3)
5. lower-triangle rest entries ~ N(0, 0.1) * scale of their row's dimension
6. Beta logits b ~ N(0, 0.5), D-2 per primitive -> beta = 4 exp(b)
7. opacity logits ~ N(0, 1)                      -> opacity = sigmoid
8. RGB ~ U[0, 1]
9. angular means ~ U[-1, 1]^3 (the reference initialiser,
   scene/beta_model.py:169-171)
10. 7-D only: temporal mean ~ U[0, 1], temporal log-scale ~ N(-2, 0.5)
    (overrides scale 6), temporal Beta logit ~ N(0, 0.5) (overrides b[:, 4]).

Leaf parameters are stored pre-activation exactly as ``BetaModel`` stores
them (scene/beta_model.py:36-52): ``scale_raw = softplus^-1(scale)``,
``opacity_raw``, ``beta_raw`` and the rest entries of ``l_triangle`` (its
first three entries are the skew rotation parameters of the model; here the
rotation comes from the quaternion instead and those entries are zero).

Cameras: position = radius * normalise(N(0, I_3)) from a second generator
(seed + 1000), looking at the origin with OpenCV axes (x right, y down,
z forward), world up +z (+y when degenerate); fx = fy = W / (2 tan 30deg),
cx = W/2, cy = H/2 (scene/beta_model.py:664-673).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np


def rest_pattern(D: int):
    """rest_i / rest_j of scene/beta_model.py:69-73: strictly-lower entries
    (torch.tril_indices order) whose row or column is >= 3."""
    ri, rj = [], []
    for r in range(D):
        for c in range(r):
            if r >= 3 or c >= 3:
                ri.append(r)
                rj.append(c)
    return np.asarray(ri, np.int32), np.asarray(rj, np.int32)


def quat_to_rotmat_np(q: np.ndarray) -> np.ndarray:
    q = q / np.linalg.norm(q, axis=-1, keepdims=True)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
        2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
        2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y),
    ], axis=-1)
    return R.reshape(-1, 3, 3)


def inv_softplus(y: np.ndarray) -> np.ndarray:
    return y + np.log(-np.expm1(-y))


@dataclass
class Scene:
    D: int
    xyz: np.ndarray          # [N,3]
    mean_rest: np.ndarray    # [N,D-3]  angular (+ temporal) means
    scale_raw: np.ndarray    # [N,D]    pre-softplus
    quats: np.ndarray        # [N,4]
    rot: np.ndarray          # [N,3,3]  R(q)
    l_triangle: np.ndarray   # [N,D(D-1)/2]
    opacity_raw: np.ndarray  # [N,1]
    beta_raw: np.ndarray     # [N,D-2]
    rgb: np.ndarray          # [N,3]

    @property
    def N(self) -> int:
        return self.xyz.shape[0]

    def astype(self, dt):
        return Scene(self.D, *(getattr(self, f).astype(dt) for f in (
            "xyz", "mean_rest", "scale_raw", "quats", "rot", "l_triangle", "opacity_raw",
            "beta_raw", "rgb")))


def make_scene(N: int, D: int = 6, seed: int = 0, dtype=np.float32) -> Scene:
    assert D in (6, 7), "the benchmark scenes are 6-D (static) or 7-D (dynamic)"
    g = np.random.Generator(np.random.PCG64(seed))
    xyz = g.uniform(-1.0, 1.0, (N, 3))
    ls_xyz = g.normal(-3.5, 0.5, (N, 3))
    q = g.normal(0.0, 1.0, (N, 4))
    q /= np.linalg.norm(q, axis=-1, keepdims=True)
    ls_ang = g.normal(0.0, 0.3, (N, 3))
    ri, rj = rest_pattern(D)
    z_rest = g.normal(0.0, 0.1, (N, ri.size))
    b = g.normal(0.0, 0.5, (N, D - 2))
    o_logit = g.normal(0.0, 1.0, (N, 1))
    rgb = g.uniform(0.0, 1.0, (N, 3))
    m_ang = g.uniform(-1.0, 1.0, (N, 3))
    scale = np.concatenate([np.exp(ls_xyz), np.exp(ls_ang)], axis=1)
    mean_rest = m_ang
    if D == 7:
        t_mean = g.uniform(0.0, 1.0, (N, 1))
        ls_t = g.normal(-2.0, 0.5, (N, 1))
        b_t = g.normal(0.0, 0.5, (N, 1))
        scale = np.concatenate([scale, np.exp(ls_t)], axis=1)
        mean_rest = np.concatenate([m_ang, t_mean], axis=1)
        b[:, 4:5] = b_t
    ltri = np.zeros((N, D * (D - 1) // 2))
    ltri[:, 3:] = z_rest * scale[:, ri]
    sc = Scene(D=D, xyz=xyz, mean_rest=mean_rest, scale_raw=inv_softplus(scale), quats=q,
               rot=quat_to_rotmat_np(q), l_triangle=ltri, opacity_raw=o_logit, beta_raw=b, rgb=rgb)
    return sc.astype(dtype)


def look_at_viewmat(pos: np.ndarray, target=(0.0, 0.0, 0.0)) -> np.ndarray:
    """World->camera [4,4] row-major, OpenCV axes (cameras.py:69-73 convention)."""
    pos = np.asarray(pos, np.float64)
    f = np.asarray(target, np.float64) - pos
    f /= np.linalg.norm(f)
    up = np.array([0.0, 0.0, 1.0])
    if abs(np.dot(f, up)) > 0.999:
        up = np.array([0.0, 1.0, 0.0])
    x = np.cross(f, up)
    x /= np.linalg.norm(x)
    y = np.cross(f, x)
    R = np.stack([x, y, f], axis=0)
    vm = np.eye(4)
    vm[:3, :3] = R
    vm[:3, 3] = -R @ pos
    return vm


def intrinsics(W: int, H: int, fov_deg: float = 60.0) -> np.ndarray:
    f = W / (2.0 * math.tan(math.radians(fov_deg) / 2.0))
    return np.array([[f, 0.0, W / 2.0], [0.0, f, H / 2.0], [0.0, 0.0, 1.0]])


def make_cameras(n: int, W: int, H: int, seed: int = 0, radius=4.0, dtype=np.float32):
    """n cameras on a sphere around the origin; radius may be (lo, hi)."""
    g = np.random.Generator(np.random.PCG64(seed + 1000))
    vms, Ks, ts = [], [], []
    for _ in range(n):
        d = g.normal(0.0, 1.0, 3)
        d /= np.linalg.norm(d)
        r = radius if np.isscalar(radius) else g.uniform(radius[0], radius[1])
        vms.append(look_at_viewmat(r * d))
        Ks.append(intrinsics(W, H))
        ts.append(g.uniform(0.0, 1.0))
    return (np.stack(vms).astype(dtype), np.stack(Ks).astype(dtype), np.asarray(ts, dtype))


def camera_center(viewmat: np.ndarray) -> np.ndarray:
    R = viewmat[:3, :3].astype(np.float64)
    t = viewmat[:3, 3].astype(np.float64)
    return -R.T @ t
