"""CPU checks of the step after the backward: the oracle's Adam, position
noise and MCMC relocation pinned by a library routine (torch.optim.Adam on
the CPU, fp64), closed forms and invariants; and the host logic of
gsplat.optim (learning-rate schedule, sampling) that runs before any
kernel."""
import math

import numpy as np
import pytest
import torch

from gsplat import optim as G
from oracle import ubs_oracle as O


def test_adam_oracle_matches_torch_adam_fp64():
    """The oracle's update is torch.optim.Adam's (CPU, fp64, three groups
    with their own learning rates, eps 1e-15) over six steps."""
    rng = np.random.Generator(np.random.PCG64(3))
    shapes, lrs = [(50, 3), (50, 1), (50, 15)], [1.6e-4, 0.05, 0.001]
    ps = [rng.normal(0, 1, s) for s in shapes]
    tp = [torch.tensor(p, requires_grad=True) for p in ps]
    opt = torch.optim.Adam([{"params": [t], "lr": lr} for t, lr in zip(tp, lrs)], lr=0.0, eps=1e-15)
    state = [(p.copy(), np.zeros(p.shape), np.zeros(p.shape)) for p in ps]
    for step in range(1, 7):
        gs = [rng.normal(0, 1, s) * 10.0 ** rng.integers(-6, 1) for s in shapes]
        for t, g in zip(tp, gs):
            t.grad = torch.tensor(g)
        opt.step()
        state = [O.adam_step(p, g, m, v, lr, step) for (p, m, v), g, lr in zip(state, gs, lrs)]
    for t, (p, m, v) in zip(tp, state):
        np.testing.assert_allclose(t.detach().numpy(), p, rtol=1e-13, atol=1e-15)
        st = opt.state[t]
        np.testing.assert_allclose(st["exp_avg"].numpy(), m, rtol=1e-13, atol=1e-300)
        np.testing.assert_allclose(st["exp_avg_sq"].numpy(), v, rtol=1e-13, atol=1e-300)


def test_adam_oracle_first_step_closed_form():
    """At step 1 the bias corrections cancel: p' = p - lr g / (|g| + eps)."""
    g = np.array([1e-3, -2.0, 5e-9, 0.0, -7e-14])
    p = np.linspace(-1, 1, g.size)
    lr, eps = 0.05, 1e-15
    p1, m1, v1 = O.adam_step(p, g, np.zeros_like(g), np.zeros_like(g), lr, 1, eps=eps)
    np.testing.assert_allclose(p1, p - lr * g / (np.abs(g) + eps), rtol=1e-12, atol=1e-17)
    np.testing.assert_allclose(m1, 0.1 * g, rtol=1e-15)
    np.testing.assert_allclose(v1, 0.001 * g * g, rtol=1e-12)


def test_xyz_covariance_closed_form():
    """With a zero skew the spatial covariance is diag(softplus(s)^2); the
    general case is R diag(s^2) R^T with R = I + skew(a01, a02, a12)
    (l_triagnle_to_rotmat_fwd.cu:16-35), and symmetric positive definite."""
    rng = np.random.Generator(np.random.PCG64(5))
    N, D = 20, 6
    lt = rng.normal(0, 0.3, (N, D * (D - 1) // 2))
    sr = rng.normal(-2, 0.5, (N, D))
    s = np.logaddexp(0.0, sr)
    lt0 = lt.copy()
    lt0[:, :3] = 0
    np.testing.assert_allclose(O.xyz_covariance(lt0, sr),
                               np.einsum("ni,ij->nij", s[:, :3] ** 2, np.eye(3)), rtol=1e-14, atol=1e-18)
    S = O.xyz_covariance(lt, sr)
    a0, a1, a2 = lt[:, 0], lt[:, 1], lt[:, 2]
    for n in range(N):
        R = np.array([[1, a0[n], a1[n]], [-a0[n], 1, a2[n]], [-a1[n], -a2[n], 1]])
        np.testing.assert_allclose(S[n], R @ np.diag(s[n, :3] ** 2) @ R.T, rtol=1e-12, atol=1e-18)
        assert np.all(np.linalg.eigvalsh(S[n]) > 0)


def test_xyz_noise_closed_forms():
    """Opaque primitives do not move ((1 - o)^100 = 0 to fp64 precision at
    o -> 1); with zero skew the noise is s^2 z (1 - o)^100 noise_lr xyz_lr;
    and it is linear in the draws."""
    rng = np.random.Generator(np.random.PCG64(6))
    N, D = 16, 7
    xyz = rng.uniform(-1, 1, (N, 3))
    lt = rng.normal(0, 0.3, (N, D * (D - 1) // 2))
    sr = rng.normal(-2, 0.5, (N, D))
    z = rng.normal(0, 1, (N, 3))
    op = np.full((N, 1), 40.0)
    np.testing.assert_allclose(O.xyz_noise(xyz, lt, sr, op, z, 1.0, 1e-4), xyz, rtol=0, atol=1e-300)
    lt[:, :3] = 0
    op = rng.normal(-3, 1, (N, 1))
    o = 1 / (1 + np.exp(-op))
    s = np.logaddexp(0.0, sr[:, :3])
    want = xyz + s ** 2 * z * (1 - o) ** 100 * 0.5 * 2e-4
    np.testing.assert_allclose(O.xyz_noise(xyz, lt, sr, op, z, 0.5, 2e-4), want, rtol=1e-12)
    lt = rng.normal(0, 0.3, lt.shape)
    zero = np.zeros_like(xyz)
    a = O.xyz_noise(zero, lt, sr, op, z, 1.0, 1e-3)
    b = O.xyz_noise(zero, lt, sr, op, 3 * z, 1.0, 1e-3)
    np.testing.assert_allclose(b, 3 * a, rtol=1e-12)


def test_split_opacity_composes_to_the_original():
    """ratio + 1 copies of opacity o' let through what one of opacity o did:
    1 - (1 - o')^(ratio+1) = o, unless a clamp (0.005, 1 - 2^-23) applies."""
    rng = np.random.Generator(np.random.PCG64(7))
    raw = rng.normal(0, 2, 200)
    ratio = rng.integers(1, 6, 200)
    on = 1 / (1 + np.exp(-O.split_opacity_raw(raw, ratio)))
    o = 1 / (1 + np.exp(-raw))
    free = (on > 0.005 + 1e-12) & (on < 1 - 2.0 ** -23 - 1e-12)
    assert free.sum() > 150
    np.testing.assert_allclose(1 - (1 - on[free]) ** (ratio[free] + 1), o[free], rtol=1e-10)
    assert np.all(on[~free] >= 0.005 - 1e-12)
    assert np.allclose(O.split_opacity_raw(np.array([-20.0]), np.array([3])), math.log(0.005 / 0.995))


def test_mcmc_relocate_oracle_invariants():
    """dst rows become their src rows in every leaf but opacity; src and dst
    share the split opacity; only src rows' moments are zeroed."""
    rng = np.random.Generator(np.random.PCG64(8))
    N = 30
    leaves = {"xyz": rng.normal(0, 1, (N, 3)), "opacity_raw": rng.normal(0, 1, (N, 1)),
              "l_triangle": rng.normal(0, 1, (N, 15))}
    m = {k: rng.normal(0, 1, v.shape) for k, v in leaves.items()}
    v = {k: rng.uniform(0, 1, v.shape) for k, v in leaves.items()}
    src = np.array([3, 3, 7, 11])
    dst = np.array([0, 1, 2, 29])
    ratio = np.array([2, 2, 1, 1])
    L, M, V = O.mcmc_relocate(leaves, m, v, src, dst, ratio)
    for k in ("xyz", "l_triangle"):
        np.testing.assert_array_equal(L[k][dst], leaves[k][src])
    np.testing.assert_array_equal(L["opacity_raw"][dst], L["opacity_raw"][src])
    o3 = 1 / (1 + np.exp(-leaves["opacity_raw"][3, 0]))
    on3 = 1 / (1 + np.exp(-L["opacity_raw"][3, 0]))
    assert (1 - (1 - on3) ** 3) == pytest.approx(o3, rel=1e-12)
    untouched = np.setdiff1d(np.arange(N), np.concatenate([src, dst]))
    for k in leaves:
        np.testing.assert_array_equal(M[k][src], 0)
        np.testing.assert_array_equal(V[k][src], 0)
        np.testing.assert_array_equal(M[k][dst], m[k][dst])
        np.testing.assert_array_equal(L[k][untouched], leaves[k][untouched])


def test_expon_lr_schedule():
    """get_expon_lr_func: lr_init at 0, lr_final at max_steps and beyond,
    geometric mean half-way, 0 for negative steps; the delay multiplies by
    lr_delay_mult at step 0 and by 1 from lr_delay_steps on."""
    f = G.expon_lr(1.6e-4, 1.6e-6, lr_delay_mult=0.01, max_steps=30_000)
    assert f(0) == pytest.approx(1.6e-4, rel=1e-14)
    assert f(30_000) == pytest.approx(1.6e-6, rel=1e-12)
    assert f(90_000) == pytest.approx(1.6e-6, rel=1e-12)
    assert f(15_000) == pytest.approx(math.sqrt(1.6e-4 * 1.6e-6), rel=1e-12)
    assert f(-1) == 0.0
    d = G.expon_lr(1.0, 1e-2, lr_delay_steps=100, lr_delay_mult=0.1, max_steps=1000)
    assert d(0) == pytest.approx(0.1, rel=1e-14)
    assert d(100) == pytest.approx(1e-2 ** 0.1, rel=1e-12)


def test_sample_alives_ratio_is_draw_count():
    """_sample_alives: draws map through alive_indices, and ratio is how
    often each drawn index was drawn; zero-probability rows are never drawn."""
    gen = torch.Generator().manual_seed(0)
    probs = torch.tensor([0.0, 1.0, 3.0, 0.0, 2.0])
    alive = torch.tensor([10, 11, 12, 13, 14])
    idx, ratio = G.sample_alives(probs, 64, alive, gen)
    assert idx.shape == (64,) and ratio.shape == (64,)
    assert not bool(((idx == 10) | (idx == 13)).any())
    counts = {int(i): int((idx == i).sum()) for i in idx.unique()}
    assert all(int(r) == counts[int(i)] for i, r in zip(idx, ratio))


def test_leaf_groups_cover_the_reference_groups():
    """The seven groups of training_setup (beta_model.py:240-267), each a
    BetaParams field, and OptimizationParams' defaults those of
    arguments/__init__.py:81-104."""
    assert [g for g, _ in G.LEAF_GROUPS] == ["xyz", "mean", "rgb", "opacity", "beta", "scale", "l_triangle"]
    from gsplat.beta import BetaParams
    fields = set(BetaParams.__dataclass_fields__)
    assert all(f in fields for _, f in G.LEAF_GROUPS)
    o = G.OptimizationParams()
    assert (o.position_lr_init, o.position_lr_final, o.rgb_lr, o.mean_lr, o.opacity_lr, o.beta_lr, o.scale_lr,
            o.l_triangle_lr, o.noise_lr) == (0.00016, 0.0000016, 0.001, 0.001, 0.05, 0.001, 0.005, 0.001, 1.0)
