"""The projection's FP64 branch for near-singular 2-D covariances, on the GPU.

The reference inverts the blurred 2-D covariance in fp32
(fully_fused_projection_fwd.cu:137-150, utils.cuh:438-448).  Where
det < 2^-8 c00 c11 this build takes the determinant, the culling test, the
inverse and the compensation in FP64 (the oracle mirrors it and is pinned by
the exact inverse in tests/test_oracle_math.py).  Here:

* constructed 2-D covariances with the exact inverse known: the GPU culls by
  the exact determinant's sign and its conic is the exact inverse to 1 ulp
  (bit-identical to the oracle);
* conditioned 6-D primitives with strong spatial/view cross-covariance and
  unequal Beta factors below 4, whose mirrored conditioned covariance is
  often indefinite: the drop-in chain (conditioning, upper-triangle gather,
  projection; the fused Beta projection runs the same per-primitive code) is
  bit-exact against the oracle's float32 restatement, on hundreds of
  primitives that take the branch and hundreds that it culls.
"""
import numpy as np
import pytest
import torch

from gsplat import scenes
from gsplat.cuda import _wrapper as W
from oracle import ubs_oracle as O

from util import (assert_same, blurred_det_ratio, cross_correlated_6d, expected_conic_2d,
                  near_singular_2d_cases)

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def test_near_singular_inverse_exact(device):
    means, cov6, vm, K, Wd, Hd, cs, c01s = near_singular_2d_cases()
    print(f"near-singular 2-D cases: {cs.size}", flush=True)
    r, m2, d, cn, _ = W.fully_fused_projection(cu(means), cu(cov6), None, None, cu(vm), cu(K), Wd, Hd)
    r, m2, cn = host(r), host(m2), host(cn)
    wr, wm2, wd, wcn, _ = O.proj_fwd(means, cov6, None, None, vm, K, Wd, Hd)
    assert_same(r, wr, "radii")
    vis = wr > 0
    assert_same(cn[vis], wcn[vis], "conics")
    assert_same(m2[vis], wm2[vis], "means2d")
    n_branch = 0
    for i in range(cs.size):
        det_exact, exact, _, fp32, branch = expected_conic_2d(cs[i], c01s[i])
        n_branch += int(branch)
        if det_exact <= 0:
            assert r[0, i] == 0, i
        elif branch:
            assert (np.abs(cn[0, i].astype(np.float64) - exact) <= np.spacing(np.abs(exact))).all(), i
        else:
            assert (cn[0, i] == fp32).all(), i
    assert n_branch >= 20


@pytest.mark.parametrize("seed", [3, 4])
def test_conditioned_indefinite_chain_bit_exact(device, seed):
    N = 100_000
    m6, cov, op, be, q = cross_correlated_6d(N, seed=seed)
    vms, Ks, _ = scenes.make_cameras(2, 256, 256, seed=seed)
    print(f"cross-correlated 6-D: N={N} seed={seed}", flush=True)
    om, oV, oo = O.cond_fwd(m6, cov, op, be, q)
    m, V, o = W.cond_mean_convariance_opacity(cu(m6), cu(cov), cu(op), cu(be), cu(q))
    assert_same(host(m), om, "m_cond")
    assert_same(host(V), oV, "V_cond")
    r, m2, d, cn, _ = W.fully_fused_projection_covars33(m, V, cu(vms), cu(Ks), 256, 256)
    tri = ([0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2])
    wr, wm2, wd, wcn, _ = O.proj_fwd(om, np.ascontiguousarray(oV[:, tri[0], tri[1]]), None, None, vms, Ks, 256, 256)
    r, m2, d, cn = host(r), host(m2), host(d), host(cn)
    assert_same(r, wr, "radii")
    vis = wr > 0
    for got, want, k in ((m2, wm2, "means2d"), (d, wd, "depths"), (cn, wcn, "conics")):
        assert_same(got[vis], want[vis], k)
    n_near_vis = 0
    for c in range(2):
        rat = blurred_det_ratio(om, oV, vms[c], Ks[c], 256, 256)
        near = (rat > 0) & (rat < 2.0 ** -8)
        indef = rat < -1e-3
        n_near_vis += int((r[c][near] > 0).sum())
        print(f"camera {c}: {int(near.sum())} near-singular ({int((r[c][near] > 0).sum())} on screen), "
              f"{int(indef.sum())} indefinite", flush=True)
        assert indef.sum() >= 100 and (r[c][indef] == 0).all()
    assert n_near_vis >= 50
