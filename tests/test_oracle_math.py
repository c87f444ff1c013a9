"""Pins the oracle's covariance, conditioning and projection against
mathematics that does not reuse its code: dense numpy / scipy linear algebra,
closed forms, torch FP64 autograd of dense formulations and central finite
differences.  CPU only."""
import numpy as np
import pytest
import scipy.linalg
import torch

from gsplat import scenes
from oracle import ubs_oracle as O



@pytest.fixture(autouse=True)
def _float64_default():
    """Dense torch references here run in FP64; restore the process default
    afterwards (a leaked float64 default makes later tests create float64
    tensors the fp32 kernels reject)."""
    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(old)


def dense_L(rot, s, ltri, ri, rj):
    N, D = s.shape
    L = np.zeros((N, D, D))
    L[:, :3, :3] = rot * s[:, None, :3]
    for k in range(3, D):
        L[:, k, k] = s[:, k]
    L[:, ri, rj] = ltri[:, 3:3 + ri.size]
    return L


def rand_covar_inputs(N, D, seed):
    g = np.random.Generator(np.random.PCG64(seed))
    rot = scenes.quat_to_rotmat_np(g.normal(0, 1, (N, 4)))
    s = np.exp(g.normal(-0.5, 0.4, (N, D)))
    lt = g.normal(0, 0.3, (N, D * (D - 1) // 2))
    ri, rj = scenes.rest_pattern(D)
    return rot, s, lt, ri, rj


# ---------------------------------------------------------------- rotmat
def test_rotmat_closed_form():
    l = np.random.Generator(np.random.PCG64(0)).normal(0, 1, (50, 3))
    R = O.rotmat_fwd(l)
    A = np.zeros((50, 3, 3))
    A[:, 0, 1], A[:, 0, 2], A[:, 1, 2] = l[:, 0], l[:, 1], l[:, 2]
    np.testing.assert_array_equal(R, np.eye(3) + A - A.transpose(0, 2, 1))
    G = np.random.Generator(np.random.PCG64(1)).normal(0, 1, (50, 3, 3))
    # d<G, R(l)>/dl = <G, dR/dl>; dR/dl0 has +1 at (0,1), -1 at (1,0)
    np.testing.assert_allclose(O.rotmat_bwd(G), np.stack(
        [G[:, 0, 1] - G[:, 1, 0], G[:, 0, 2] - G[:, 2, 0], G[:, 1, 2] - G[:, 2, 1]], 1))


# ---------------------------------------------------------------- covariance
@pytest.mark.parametrize("D", [4, 6, 7, 9])
def test_covar_matches_dense_LLt(D):
    rot, s, lt, ri, rj = rand_covar_inputs(200, D, D)
    L = dense_L(rot, s, lt, ri, rj)
    want = L @ L.transpose(0, 2, 1)
    got = O.covar_fwd(rot, s, lt, ri, rj)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    assert np.allclose(got, got.transpose(0, 2, 1))
    assert (np.linalg.eigvalsh(got) > 0).all()
    sp = O.covar_fwd(rot, s, lt, ri, rj, spatial_block=True)
    M = rot * s[:, None, :3]
    np.testing.assert_allclose(sp, M @ M.transpose(0, 2, 1), rtol=1e-12)
    # float32 instantiation agrees to rounding
    got32 = O.covar_fwd(*(a.astype(np.float32) for a in (rot, s, lt)), ri, rj)
    np.testing.assert_allclose(got32, want, rtol=2e-5, atol=1e-6)


@pytest.mark.parametrize("D,spatial", [(6, False), (7, False), (6, True)])
def test_covar_bwd_matches_autograd(D, spatial):
    rot, s, lt, ri, rj = rand_covar_inputs(64, D, 10 + D)
    G = np.random.Generator(np.random.PCG64(3)).normal(0, 1, (64, 3, 3) if spatial else (64, D, D))
    tr, ts, tl = (torch.tensor(a, requires_grad=True) for a in (rot, s, lt))
    Lxyz = tr * ts[:, None, :3]
    if spatial:
        out = Lxyz @ Lxyz.transpose(1, 2)
    else:
        L = torch.zeros(64, D, D)
        L[:, :3, :3] = Lxyz
        idx = torch.arange(3, D)
        L[:, idx, idx] = ts[:, 3:]
        L[:, torch.tensor(ri).long(), torch.tensor(rj).long()] = tl[:, 3:]
        out = L @ L.transpose(1, 2)
    (out * torch.tensor(G)).sum().backward()
    vr, vs, vl = O.covar_bwd(rot, s, lt, ri, rj, spatial, G)
    np.testing.assert_allclose(vr, tr.grad.numpy(), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(vs, ts.grad.numpy(), rtol=1e-10, atol=1e-12)
    if spatial:
        assert tl.grad is None and (vl == 0).all()
    else:
        np.testing.assert_allclose(vl, tl.grad.numpy(), rtol=1e-10, atol=1e-12)
    assert (vl[:, :3] == 0).all()


# ---------------------------------------------------------------- conditioning
def cond_inputs(N, D, seed, beta_lo=0.5, beta_hi=8.0):
    rot, s, lt, ri, rj = rand_covar_inputs(N, D, seed)
    V = O.covar_fwd(rot, s, lt, ri, rj)
    g = np.random.Generator(np.random.PCG64(seed + 100))
    m = g.normal(0, 1, (N, D))
    o = g.uniform(0.1, 0.9, (N, 1))
    b = g.uniform(beta_lo, beta_hi, (N, D - 3))
    q = m[:, 3:] + g.normal(0, 0.7, (N, D - 3))
    return m, V, o, b, q


@pytest.mark.parametrize("D", [4, 6, 7])
def test_cond_is_gaussian_conditional_when_beta_ge_4(D):
    m, V, o, b, q = cond_inputs(100, D, 5, beta_lo=4.0, beta_hi=9.0)
    om, oV, oo = O.cond_fwd(m, V, o, b, q)
    for i in range(100):
        V11, V12, V21, V22 = V[i, :3, :3], V[i, :3, 3:], V[i, 3:, :3], V[i, 3:, 3:]
        np.testing.assert_allclose(om[i], m[i, :3] + V12 @ np.linalg.solve(V22, q[i] - m[i, 3:]), rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(oV[i], V11 - V12 @ np.linalg.solve(V22, V21), rtol=1e-9, atol=1e-12)
        y = scipy.linalg.solve_triangular(np.linalg.cholesky(V22), q[i] - m[i, 3:], lower=True)
        want = o[i, 0] * np.prod((1 - np.tanh(y ** 2)) ** b[i])
        if want > 1e-30:  # (1 - tanh(y^2)) underflows differently for huge y
            np.testing.assert_allclose(oo[i, 0], want, rtol=1e-8)


def test_cond_special_cases():
    D = 6
    m, V, o, b, q = cond_inputs(50, D, 8)
    V0 = V.copy()
    V0[:, :3, 3:] = 0
    V0[:, 3:, :3] = 0
    om, oV, oo = O.cond_fwd(m, V0, o, b, q)
    np.testing.assert_array_equal(om, m[:, :3])  # no cross covariance: nothing moves
    np.testing.assert_array_equal(oV, V0[:, :3, :3])
    om, oV, oo = O.cond_fwd(m, V, o, b, np.ascontiguousarray(m[:, 3:]))
    np.testing.assert_array_equal(oo, o)  # query at the mean keeps the opacity
    np.testing.assert_array_equal(om, m[:, :3])
    # beta < 4 scales the regression by beta/4 per dimension
    b2 = np.full_like(b, 2.0)
    om2, oV2, _ = O.cond_fwd(m, V, o, b2, q)
    for i in range(5):
        r = V[i, :3, 3:] @ np.linalg.inv(V[i, 3:, 3:]) * 0.5
        np.testing.assert_allclose(om2[i], m[i, :3] + r @ (q[i] - m[i, 3:]), rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(oV2[i], V[i, :3, :3] - r @ V[i, 3:, :3], rtol=1e-9, atol=1e-12)


def torch_cond(m, V, o, b, q):
    """Dense torch formulation (inverse / cholesky / triangular solve)."""
    V11, V12, V21, V22 = V[:, :3, :3], V[:, :3, 3:], V[:, 3:, :3], V[:, 3:, 3:]
    x = q - m[:, 3:]
    badj = torch.clamp(b / 4.0, max=1.0)
    rb = (V12 @ torch.linalg.inv(V22)) * badj[:, None, :]
    om = m[:, :3] + (rb @ x[..., None])[..., 0]
    oV = V11 - rb @ V21
    y = torch.linalg.solve_triangular(torch.linalg.cholesky(V22), x[..., None], upper=False)[..., 0]
    d = torch.tanh(y * y).clamp(0.0, 1.0 - np.finfo(np.float64).eps)
    oo = o * torch.prod((1 - d) ** b, dim=-1, keepdim=True)
    return om, oV, oo


@pytest.mark.parametrize("D", [6, 7])
def test_cond_bwd_matches_autograd_and_fd(D):
    """The exact mode is the true gradient (autograd and finite differences)."""
    m, V, o, b, q = cond_inputs(40, D, 20 + D)
    g = np.random.Generator(np.random.PCG64(2))
    vm, vV, vo = g.normal(0, 1, (40, 3)), g.normal(0, 1, (40, 3, 3)), g.normal(0, 1, (40, 1))
    tm, tV, to, tb = (torch.tensor(a, requires_grad=True) for a in (m, V, o, b))
    om, oV, oo = torch_cond(tm, tV, to, tb, torch.tensor(q))
    ((om * torch.tensor(vm)).sum() + (oV * torch.tensor(vV)).sum() + (oo * torch.tensor(vo)).sum()).backward()
    gm, gV, go, gb = O.cond_bwd(m, V, o, b, q, vm, vV, vo, exact=True)
    np.testing.assert_allclose(gm, tm.grad.numpy(), rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(go, to.grad.numpy(), rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(gb, tb.grad.numpy(), rtol=1e-8, atol=1e-10)
    # covariance gradients agree on every symmetric perturbation direction
    # (the two formulations may split a gradient differently between (i,j)
    # and (j,i) where the function only sees their sum)
    sym = lambda a: a + np.swapaxes(a, 1, 2)
    np.testing.assert_allclose(sym(gV), sym(tV.grad.numpy()), rtol=1e-7, atol=1e-9)
    # central finite differences of the oracle itself on a symmetric direction
    dV = g.normal(0, 1, V.shape)
    dV = 0.5 * (dV + np.swapaxes(dV, 1, 2))
    eps = 1e-6

    def f(Vx):
        a, B, c = O.cond_fwd(m, Vx, o, b, q)
        return (a * vm).sum() + (B * vV).sum() + (c * vo).sum()

    fd = (f(V + eps * dV) - f(V - eps * dV)) / (2 * eps)
    np.testing.assert_allclose((gV * dV).sum(), fd, rtol=1e-5)


@pytest.mark.parametrize("D", [6, 7])
def test_cond_bwd_reference_mode_is_the_transposed_solve(D):
    """Default mode reproduces the reference, whose Cholesky backward forms
    G = L^-1 S^T L^-1 instead of L^-T S L^-1 (S = Phi(L^T dL)).  Pinned with
    torch autograd for dL (gradient w.r.t. the Cholesky factor) and numpy
    solves for the rest; only the query-block opacity path differs."""
    m, V, o, b, q = cond_inputs(30, D, 40 + D)
    g = np.random.Generator(np.random.PCG64(6))
    vo = g.normal(0, 1, (30, 1))
    zero_m, zero_V = np.zeros((30, 3)), np.zeros((30, 3, 3))
    ref = O.cond_bwd(m, V, o, b, q, zero_m, zero_V, vo)
    exact = O.cond_bwd(m, V, o, b, q, zero_m, zero_V, vo, exact=True)
    for k in (0, 2, 3):  # means, opacities, betas agree
        np.testing.assert_array_equal(ref[k], exact[k])
    np.testing.assert_array_equal(ref[1][:, :3, :], exact[1][:, :3, :])
    C = D - 3
    V22 = V[:, 3:, 3:]
    L = np.linalg.cholesky(V22)
    tL = torch.tensor(L, requires_grad=True)
    x = torch.tensor(q - m[:, 3:])
    y = torch.linalg.solve_triangular(tL, x[..., None], upper=False)[..., 0]
    d = torch.tanh(y * y)
    oo = torch.tensor(o) * torch.prod((1 - d) ** torch.tensor(b), -1, keepdim=True)
    (oo * torch.tensor(vo)).sum().backward()
    gL = np.tril(tL.grad.numpy())
    X = np.tril(np.swapaxes(L, 1, 2) @ gL)
    X[:, range(C), range(C)] *= 0.5
    Linv = np.linalg.inv(L)
    G_ref = Linv @ np.swapaxes(X, 1, 2) @ Linv
    G_true = np.swapaxes(Linv, 1, 2) @ X @ Linv
    np.testing.assert_allclose(ref[1][:, 3:, 3:], 0.5 * (G_ref + np.swapaxes(G_ref, 1, 2)), rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(exact[1][:, 3:, 3:], 0.5 * (G_true + np.swapaxes(G_true, 1, 2)), rtol=1e-8,
                               atol=1e-12)
    assert np.abs(ref[1] - exact[1]).max() > 1e-3 * np.abs(exact[1]).max()


# ---------------------------------------------------------------- projection
def cam(W=640, H=480, pos=(0.3, -0.2, -4.0), target=(0, 0, 0)):
    vm = scenes.look_at_viewmat(np.asarray(pos), target)
    return vm[None], scenes.intrinsics(W, H)[None]


def test_projection_on_axis_closed_form():
    W, H = 640, 480
    vm = np.eye(4)[None]
    K = scenes.intrinsics(W, H)[None]
    f = K[0, 0, 0]
    for z, s in ((4.0, 0.05), (2.0, 0.3), (10.0, 0.01)):
        means = np.array([[0.0, 0.0, z]])
        cov6 = np.array([[s * s, 0, 0, s * s, 0, s * s]])
        r, m2, d, cn, comp = O.proj_fwd(means, cov6, None, None, vm, K, W, H, calc_compensations=True)
        v = f * f * s * s / (z * z)
        np.testing.assert_allclose(m2[0, 0], [W / 2, H / 2])
        np.testing.assert_allclose(cn[0, 0], [1 / (v + 0.3), 0, 1 / (v + 0.3)], rtol=1e-12)
        assert r[0, 0] == int(np.ceil(np.sqrt(v + 0.3 + np.sqrt(max(0.01, 0.0)))))
        np.testing.assert_allclose(comp[0, 0], np.sqrt(v * v / (v + 0.3) ** 2), rtol=1e-12)
        assert d[0, 0] == z


def test_projection_culling():
    W, H = 200, 100
    vm = np.eye(4)[None]
    K = scenes.intrinsics(W, H)[None]
    cov6 = np.tile([[1e-4, 0, 0, 1e-4, 0, 1e-4]], (5, 1))
    means = np.array([[0, 0, 0.005], [0, 0, 5.0], [0, 0, 2e10], [50.0, 0, 1.0], [0, 0, 3.0]])
    r, *_ = O.proj_fwd(means, cov6, None, None, vm, K, W, H)
    assert list(r[0]) == [0, r[0, 1], 0, 0, r[0, 4]] and r[0, 1] > 0 and r[0, 4] > 0
    r2, *_ = O.proj_fwd(means, cov6, None, None, vm, K, W, H, radius_clip=float(r[0, 1]))
    assert r2[0, 1] == 0  # radius <= radius_clip is culled


def torch_project(means, cov6, vm, K, W, H, eps2d=0.3):
    """Dense torch EWA projection (same clamp), for autograd."""
    R, t = vm[:, :3, :3], vm[:, :3, 3]
    S = torch.stack([cov6[:, 0], cov6[:, 1], cov6[:, 2], cov6[:, 1], cov6[:, 3], cov6[:, 4], cov6[:, 2], cov6[:, 4],
                     cov6[:, 5]], -1).reshape(-1, 3, 3)
    pc = torch.einsum("cij,nj->cni", R, means) + t[:, None]
    Sc = R[:, None] @ S[None] @ R[:, None].transpose(-1, -2)
    fx, fy, cx, cy = K[:, 0, 0, None], K[:, 1, 1, None], K[:, 0, 2, None], K[:, 1, 2, None]
    x, y, z = pc[..., 0], pc[..., 1], pc[..., 2]
    tfx, tfy = 0.5 * W / fx, 0.5 * H / fy
    tx = z * torch.minimum((W - cx) / fx + 0.3 * tfx, torch.maximum(-(cx / fx + 0.3 * tfx), x / z))
    ty = z * torch.minimum((H - cy) / fy + 0.3 * tfy, torch.maximum(-(cy / fy + 0.3 * tfy), y / z))
    zero = torch.zeros_like(z)
    J = torch.stack([fx / z, zero, -fx * tx / z ** 2, zero, fy / z, -fy * ty / z ** 2], -1).reshape(*z.shape, 2, 3)
    c2 = J @ Sc @ J.transpose(-1, -2) + eps2d * torch.eye(2)
    det = c2[..., 0, 0] * c2[..., 1, 1] - c2[..., 0, 1] * c2[..., 1, 0]
    conic = torch.stack([c2[..., 1, 1] / det, -c2[..., 0, 1] / det, c2[..., 0, 0] / det], -1)
    m2 = torch.stack([fx * x / z + cx, fy * y / z + cy], -1)
    return m2, z, conic


def test_projection_bwd_matches_autograd():
    g = np.random.Generator(np.random.PCG64(4))
    N = 60
    means = g.uniform(-1, 1, (N, 3))
    M = scenes.quat_to_rotmat_np(g.normal(0, 1, (N, 4))) * g.uniform(0.02, 0.3, (N, 1, 3))
    C = M @ M.transpose(0, 2, 1)
    cov6 = C[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
    vms = np.stack([scenes.look_at_viewmat(p) for p in ([0.5, 0.2, -4.0], [3.0, 1.0, 2.0])])
    Ks = np.stack([scenes.intrinsics(320, 240)] * 2)
    r, m2, d, cn, _ = O.proj_fwd(means, cov6, None, None, vms, Ks, 320, 240)
    vis = r > 0
    assert vis.sum() > 60
    vm2, vd, vcn = g.normal(0, 1, m2.shape) * vis[..., None], g.normal(0, 1, d.shape) * vis, g.normal(0, 1, cn.shape) * vis[..., None]
    tmean, tcov, tvm = (torch.tensor(a, requires_grad=True) for a in (means, cov6, vms))
    tm2, tz, tcn = torch_project(tmean, tcov, tvm, torch.tensor(Ks), 320, 240)
    np.testing.assert_allclose(tm2.detach().numpy()[vis], m2[vis], rtol=1e-10)
    np.testing.assert_allclose(tcn.detach().numpy()[vis], cn[vis], rtol=1e-9)
    ((tm2 * torch.tensor(vm2)).sum() + (tz * torch.tensor(vd)).sum() + (tcn * torch.tensor(vcn)).sum()).backward()
    v_means, v_cov, _, _, v_vm = O.proj_bwd(means, cov6, None, None, vms, Ks, 320, 240, 0.3, r, cn, None, vm2, vd, vcn,
                                            viewmats_requires_grad=True)
    np.testing.assert_allclose(v_means, tmean.grad.numpy(), rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(v_cov, tcov.grad.numpy(), rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(v_vm[:, :3, :], tvm.grad.numpy()[:, :3, :], rtol=1e-7, atol=1e-8)


def test_projection_quats_path_matches_covariance_path():
    g = np.random.Generator(np.random.PCG64(5))
    N = 100
    q = g.normal(0, 1, (N, 4))
    s = g.uniform(0.02, 0.2, (N, 3))
    M = scenes.quat_to_rotmat_np(q) * s[:, None, :]
    C = M @ M.transpose(0, 2, 1)
    cov6 = C[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
    means = g.uniform(-1, 1, (N, 3))
    vm, K = cam()
    a = O.proj_fwd(means, cov6, None, None, vm, K, 640, 480)
    b = O.proj_fwd(means, None, q, s, vm, K, 640, 480)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_allclose(a[3], b[3], rtol=1e-9)


# ------------------------------------------------------------ image loss oracle
def test_loss_window_is_the_reference_gaussian():
    g = O.gauss_window_f32()
    assert g.dtype == np.float32 and g.shape == (11,)
    assert abs(float(g.sum(dtype=np.float64)) - 1.0) < 1e-6 and np.allclose(g, g[::-1])
    assert np.isclose(g[5] / g[4], np.exp(1.0 / 4.5), rtol=1e-6)  # exp(-(k-5)^2 / (2 * 1.5^2))


def test_ssim_map_identities():
    rng = np.random.Generator(np.random.PCG64(0))
    x = rng.uniform(0, 1, (1, 20, 24, 3))
    y = rng.uniform(0, 1, (1, 20, 24, 3))
    assert np.allclose(O.ssim_map(x, x), 1.0, atol=1e-12)          # S(x, x) = 1
    assert np.allclose(O.ssim_map(x, y), O.ssim_map(y, x), atol=1e-12)  # symmetric
    # constant images a, b away from the zero-padded border: with m the 2-D
    # window's mass (1 up to float32 rounding), mu = a m, b m, E[x^2] = a^2 m,
    # E[xy] = a b m, so the variances are a^2 (m - m^2) etc.; for m = 1 this is
    # S = (2ab + C1) / (a^2 + b^2 + C1)
    a, b = 0.3, 0.7
    s = O.ssim_map(np.full((1, 30, 30, 1), a), np.full((1, 30, 30, 1), b))
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    m = float(O.gauss_window_f32().sum(dtype=np.float64)) ** 2
    v = m - m * m
    want = ((2 * a * b * m * m + C1) * (2 * a * b * v + C2)) / ((a * a + b * b) * m * m + C1) / ((a * a + b * b) * v + C2)
    assert np.allclose(s[0, 5:-5, 5:-5, 0], want, rtol=1e-9)
    assert abs(want - (2 * a * b + C1) / (a * a + b * b + C1)) < 1e-4
    assert not np.allclose(s[0, 0, 0, 0], s[0, 15, 15, 0])  # the border sees the zero padding


@pytest.mark.parametrize("lam", [0.0, 0.2, 1.0])
def test_image_loss_gradient_finite_differences(lam):
    rng = np.random.Generator(np.random.PCG64(1))
    x = rng.uniform(0, 1, (1, 9, 8, 2))
    y = rng.uniform(0, 1, (1, 9, 8, 2))
    loss, grad = O.image_loss(x, y, lam)
    eps = 1e-6
    for idx in [(0, 0, 0, 0), (0, 4, 3, 1), (0, 8, 7, 0), (0, 2, 6, 1)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += eps
        xm[idx] -= eps
        fd = (O.image_loss(xp, y, lam)[0] - O.image_loss(xm, y, lam)[0]) / (2 * eps)
        assert abs(fd - grad[idx]) < 1e-7 + 1e-5 * abs(fd), (idx, fd, grad[idx])
    # L1 only: the mean of |x - y|, gradient sign / n
    l1, g1 = O.image_loss(x, y, 0.0)
    assert np.isclose(l1, np.abs(x - y).mean()) and np.allclose(g1, np.sign(x - y) / x.size)


def test_projection_near_singular_inverse_is_fp64():
    """Blurred 2-D covariances from near-singular to indefinite, with the exact
    inverse known (util.near_singular_2d_cases): where det < 2^-8 c00 c11 the
    projection culls by the sign of the exact determinant and its conic is
    the exact inverse to 1 ulp; elsewhere it is the fp32 formula bit for bit.
    The fp32 formula alone is far off on the near-singular cases (a
    cancellation of up to 24 bits), which is what the FP64 branch is for."""
    from util import expected_conic_2d, near_singular_2d_cases

    means, cov6, vm, K, W, H, cs, c01s = near_singular_2d_cases()
    r, m2, d, cn, _ = O.proj_fwd(means, cov6, None, None, vm, K, W, H)
    n_branch = n_fp32_bad = 0
    for i in range(cs.size):
        det_exact, exact, det32, fp32, branch = expected_conic_2d(cs[i], c01s[i])
        if branch:
            n_branch += 1
            if det_exact <= 0:
                assert r[0, i] == 0, (i, cs[i], c01s[i])
                n_fp32_bad += int(det32 > 0)
                continue
            assert r[0, i] > 0, i
            ulp = np.spacing(np.abs(exact))
            assert (np.abs(cn[0, i].astype(np.float64) - exact) <= ulp).all(), (i, cn[0, i], exact)
            if fp32 is None or np.abs(fp32.astype(np.float64) / exact - 1.0).max() > 1e-3:
                n_fp32_bad += 1
        else:
            assert det_exact > 0 and r[0, i] > 0
            assert (cn[0, i] == fp32).all(), (i, cn[0, i], fp32)
    assert n_branch >= 20 and n_branch < cs.size
    assert n_fp32_bad >= 10  # the fp32 formula mis-culls or errs by > 1e-3 there
