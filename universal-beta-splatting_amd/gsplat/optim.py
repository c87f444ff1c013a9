"""The step after the backward: Adam over the scene's leaves and the MCMC
relocation / addition / noise of Beta primitives, on the GPU.

``BetaAdam`` is the optimiser BetaModel.training_setup builds
(scene/beta_model.py:239-284): one Adam group per leaf with the reference's
learning rates, eps 1e-15, the exponential position schedule of
utils/general_utils.py:34-67, moments kept per leaf, all leaves updated by
one HIP launch (``ubs_adam_step``).  ``relocate_gs`` / ``add_new_gs`` /
``add_xyz_noise`` are train.py:150-168 and beta_model.py:553-658: the
sampling stays in torch (``torch.multinomial`` as the reference draws it);
the row copies, opacity split and moment resets are one HIP kernel pair
(``ubs_mcmc_relocate``), the covariance-shaped noise one kernel
(``ubs_xyz_noise``).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import torch
from torch import Tensor

from .beta import BetaParams
from .cuda._backend import _F32, _I64, lib, ptr, require_cuda, require_dtype, stream_of
from .cuda._wrapper import _call

# Leaf order of BetaModel's optimiser groups (beta_model.py:240-267) -> BetaParams field
LEAF_GROUPS: Tuple[Tuple[str, str], ...] = (
    ("xyz", "xyz"), ("mean", "mean_rest"), ("rgb", "rgb"), ("opacity", "opacity_raw"),
    ("beta", "beta_raw"), ("scale", "scale_raw"), ("l_triangle", "l_triangle"),
)


@dataclass
class OptimizationParams:
    """The learning-rate and MCMC settings of arguments/__init__.py:81-104."""
    position_lr_init: float = 0.00016
    position_lr_final: float = 0.0000016
    position_lr_delay_mult: float = 0.01
    position_lr_max_steps: int = 30_000
    rgb_lr: float = 0.001
    mean_lr: float = 0.001
    opacity_lr: float = 0.05
    beta_lr: float = 0.001
    scale_lr: float = 0.005
    l_triangle_lr: float = 0.001
    quats_lr: float = 0.001  # rotation quaternions of the benchmark scenes (no reference group)
    noise_lr: float = 1.0
    cap_max: int = 1_000_000


def expon_lr(lr_init: float, lr_final: float, lr_delay_steps: int = 0, lr_delay_mult: float = 1.0,
             max_steps: int = 1_000_000):
    """get_expon_lr_func (utils/general_utils.py:34-67): log-linear from
    lr_init at step 0 to lr_final at max_steps, with the optional
    reverse-cosine delay."""
    def helper(step: int) -> float:
        if step < 0 or (lr_init == 0.0 and lr_final == 0.0):
            return 0.0
        if lr_delay_steps > 0:
            delay = lr_delay_mult + (1 - lr_delay_mult) * math.sin(
                0.5 * math.pi * min(max(step / lr_delay_steps, 0.0), 1.0))
        else:
            delay = 1.0
        t = min(max(step / max_steps, 0.0), 1.0)
        return delay * math.exp(math.log(lr_init) * (1 - t) + math.log(lr_final) * t)
    return helper


def _ptr_array(ts: List[Optional[Tensor]]):
    return (ctypes.c_void_p * len(ts))(*[None if t is None else t.data_ptr() for t in ts])


class BetaAdam:
    """torch.optim.Adam(groups, lr=0.0, eps=1e-15) over a BetaParams' leaves,
    one fused launch per ``step()``.

    ``groups`` maps BetaParams field -> learning rate; fields absent or with
    a None tensor are not optimised.  A leaf whose ``.grad`` is None is
    skipped by a step and keeps its step count, as torch's Adam does.
    """

    def __init__(self, params: BetaParams, lrs: Dict[str, float], betas=(0.9, 0.999), eps: float = 1e-15):
        self.params = params
        self.lrs = dict(lrs)
        self.betas = (float(betas[0]), float(betas[1]))
        self.eps = float(eps)
        self.exp_avg: Dict[str, Tensor] = {}
        self.exp_avg_sq: Dict[str, Tensor] = {}
        self.steps: Dict[str, int] = {}
        self.xyz_schedule = None
        for f in self.lrs:
            t = getattr(params, f)
            if t is None:
                continue
            require_cuda(t)
            require_dtype(_F32, t)
            self.exp_avg[f] = torch.zeros_like(t, memory_format=torch.contiguous_format)
            self.exp_avg_sq[f] = torch.zeros_like(t, memory_format=torch.contiguous_format)
            self.steps[f] = 0
        if len(self.exp_avg) > 8:
            raise ValueError("at most 8 leaves per fused Adam step")

    @staticmethod
    def for_model(params: BetaParams, opt: OptimizationParams = OptimizationParams(),
                  spatial_lr_scale: float = 1.0) -> "BetaAdam":
        """The groups and learning rates of BetaModel.training_setup
        (beta_model.py:239-276); quaternions, when the scene has them, get
        ``opt.quats_lr``."""
        lrs = {"xyz": opt.position_lr_init * spatial_lr_scale, "mean_rest": opt.mean_lr, "rgb": opt.rgb_lr,
               "opacity_raw": opt.opacity_lr, "beta_raw": opt.beta_lr, "scale_raw": opt.scale_lr,
               "l_triangle": opt.l_triangle_lr}
        if params.quats is not None:
            lrs["quats"] = opt.quats_lr
        a = BetaAdam(params, lrs)
        a.xyz_schedule = expon_lr(opt.position_lr_init * spatial_lr_scale, opt.position_lr_final * spatial_lr_scale,
                                  lr_delay_mult=opt.position_lr_delay_mult, max_steps=opt.position_lr_max_steps)
        return a

    def update_learning_rate(self, iteration: int) -> float:
        """update_learning_rate (beta_model.py:278-284): the xyz group's lr."""
        if self.xyz_schedule is None:
            raise RuntimeError("no position schedule: build the optimiser with BetaAdam.for_model")
        lr = self.xyz_schedule(iteration)
        self.lrs["xyz"] = lr
        return lr

    @torch.no_grad()
    def step(self) -> None:
        fields, ps, gs, ms, vs, sizes, lrs, steps = [], [], [], [], [], [], [], []
        for f in self.exp_avg:
            p = getattr(self.params, f)
            if p.grad is None:
                continue
            g = p.grad
            require_cuda(p, g)
            require_dtype(_F32, p, g)
            if p.shape != self.exp_avg[f].shape or g.shape != p.shape:
                raise RuntimeError(f"{f}: parameter, gradient and state shapes differ")
            self.steps[f] += 1
            fields.append(f)
            ps.append(p)
            gs.append(g)
            ms.append(self.exp_avg[f])
            vs.append(self.exp_avg_sq[f])
            sizes.append(p.numel())
            lrs.append(float(self.lrs[f]))
            steps.append(self.steps[f])
        if not fields:
            return
        n = len(fields)
        _call("ubs_adam_step", n, _ptr_array(ps), _ptr_array(gs), _ptr_array(ms), _ptr_array(vs),
              (ctypes.c_int64 * n)(*sizes), (ctypes.c_double * n)(*lrs), (ctypes.c_int64 * n)(*steps),
              self.betas[0], self.betas[1], self.eps, stream_of(ps[0]))

    def zero_grad(self, set_to_none: bool = True) -> None:
        for f in self.exp_avg:
            p = getattr(self.params, f)
            if p.grad is None:
                continue
            if set_to_none:
                p.grad = None
            else:
                p.grad.zero_()

    # -- state maintenance of the MCMC step -------------------------------------------------
    def _replace(self, field: str, t: Tensor, m: Tensor, v: Tensor) -> None:
        setattr(self.params, field, t)
        self.exp_avg[field], self.exp_avg_sq[field] = m, v


@torch.no_grad()
def add_xyz_noise(params: BetaParams, xyz_lr: float, noise_lr: float = 1.0, normal: Optional[Tensor] = None,
                  generator: Optional[torch.Generator] = None) -> Tensor:
    """train.py:159-168: xyz += Sigma_xyz (randn (1 - o)^100 noise_lr xyz_lr)
    in place; returns the normal draws used."""
    xyz = params.xyz
    N, D = params.scale_raw.shape
    if normal is None:
        normal = torch.randn(xyz.shape, device=xyz.device, dtype=torch.float32, generator=generator)
    require_cuda(xyz, params.l_triangle, params.scale_raw, params.opacity_raw, normal)
    require_dtype(_F32, xyz, params.l_triangle, params.scale_raw, params.opacity_raw, normal)
    assert normal.shape == (N, 3) and xyz.shape == (N, 3), (normal.shape, xyz.shape)
    _call("ubs_xyz_noise", N, D, ptr(xyz), ptr(params.l_triangle), ptr(params.scale_raw), ptr(params.opacity_raw),
          ptr(normal), float(noise_lr * xyz_lr), stream_of(xyz))
    return normal


def sample_alives(probs: Tensor, num: int, alive_indices: Optional[Tensor] = None,
                  generator: Optional[torch.Generator] = None) -> Tuple[Tensor, Tensor]:
    """_sample_alives (beta_model.py:572-578): indices drawn with
    replacement in proportion to probs, and how often each drawn index was
    drawn."""
    probs = probs / (probs.sum() + torch.finfo(torch.float32).eps)
    idx = torch.multinomial(probs, num, replacement=True, generator=generator)
    if alive_indices is not None:
        idx = alive_indices[idx]
    ratio = torch.bincount(idx)[idx]
    return idx, ratio


def _relocate_rows(adam: Optional[BetaAdam], params: BetaParams, src: Tensor, dst: Tensor, ratio: Tensor,
                   min_opacity: float = 0.005) -> None:
    fields = [f for _, f in LEAF_GROUPS] + (["quats"] if params.quats is not None else [])
    leaves = [getattr(params, f) for f in fields]
    require_cuda(*leaves, src, dst, ratio)
    require_dtype(_F32, *leaves)
    require_dtype(_I64, src, dst, ratio)
    m = [adam.exp_avg.get(f) if adam else None for f in fields]
    v = [adam.exp_avg_sq.get(f) if adam else None for f in fields]
    widths = [t.shape[1] for t in leaves]
    K = src.numel()
    assert dst.numel() == K and ratio.numel() == K, (src.shape, dst.shape, ratio.shape)
    if K:
        n = leaves[0].shape[0]
        lo = torch.stack([src.min(), dst.min()]).cpu()
        hi = torch.stack([src.max(), dst.max()]).cpu()
        if int(lo.min()) < 0 or int(hi.max()) >= n or any(t.shape[0] != n for t in leaves):
            raise IndexError("relocation rows out of range of the leaves")
    _call("ubs_mcmc_relocate", K, len(fields), _ptr_array(leaves), (ctypes.c_int * len(fields))(*widths),
          fields.index("opacity_raw"), _ptr_array(m), _ptr_array(v), ptr(src.contiguous()), ptr(dst.contiguous()),
          ptr(ratio.contiguous()), float(min_opacity), stream_of(leaves[0]))


@torch.no_grad()
def relocate_gs(params: BetaParams, adam: Optional[BetaAdam], dead_mask: Tensor,
                generator: Optional[torch.Generator] = None) -> int:
    """relocate_gs (beta_model.py:580-621): every dead primitive takes the
    leaves of an alive one drawn by opacity, both with the split opacity;
    the drawn rows' Adam moments are zeroed.  Returns the number moved."""
    dead = dead_mask.nonzero(as_tuple=True)[0]
    if dead.numel() == 0:
        return 0
    alive = (~dead_mask).nonzero(as_tuple=True)[0]
    if alive.numel() == 0:
        return 0
    probs = torch.sigmoid(params.opacity_raw)[alive, 0]
    src, ratio = sample_alives(probs, dead.numel(), alive, generator)
    _relocate_rows(adam, params, src, dead, ratio)
    return int(dead.numel())


@torch.no_grad()
def add_new_gs(params: BetaParams, adam: Optional[BetaAdam], cap_max: int,
               generator: Optional[torch.Generator] = None) -> int:
    """add_new_gs (beta_model.py:623-658): grow by min(cap_max, 1.02 N) - N
    primitives copied from rows drawn by opacity, with the split opacity in
    both; new rows start with zero moments, the drawn rows' are zeroed."""
    n = params.opacity_raw.shape[0]
    target = min(cap_max, int(1.02 * n))
    num = max(0, target - n)
    if num <= 0:
        return 0
    probs = torch.sigmoid(params.opacity_raw).squeeze(-1)
    src, ratio = sample_alives(probs, num, None, generator)
    fields = [f for _, f in LEAF_GROUPS] + (["quats"] if params.quats is not None else [])
    for f in fields:
        t = getattr(params, f)
        grown = torch.cat([t.detach(), t.new_empty((num,) + tuple(t.shape[1:]))], 0).requires_grad_(t.requires_grad)
        if adam is not None and f in adam.exp_avg:
            z = t.new_zeros((num,) + tuple(t.shape[1:]))
            adam._replace(f, grown, torch.cat([adam.exp_avg[f], z], 0), torch.cat([adam.exp_avg_sq[f], z], 0))
        else:
            setattr(params, f, grown)
    dst = torch.arange(n, n + num, device=src.device, dtype=torch.int64)
    _relocate_rows(adam, params, src, dst, ratio)
    return num
