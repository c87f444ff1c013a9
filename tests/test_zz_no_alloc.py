"""The allocation-free sort (ubs_radix_sort_pairs_noalloc) against the
workspace form (ubs_radix_sort_pairs) through the C ABI: two calls in a
row, the second with fewer keys and another bit range, each bit-identical
to the workspace sort and to numpy's stable argsort, status word zero."""
import ctypes

import numpy as np
import pytest
import torch

from gsplat.cuda import _wrapper as W

pytestmark = pytest.mark.gpu


def _keys(n, end_bit, seed):
    g = np.random.Generator(np.random.PCG64(seed))
    hi = g.integers(0, 1 << (end_bit - 32), n, dtype=np.int64) << 32 if end_bit > 32 else 0
    lo = g.integers(0, 1 << 32, n, dtype=np.int64)
    keys = hi | lo
    keys[: n // 7] = keys[n // 7: 2 * (n // 7)]  # equal keys: stability matters
    return keys, g.permutation(n).astype(np.int32)


def _noalloc(keys, vals, begin_bit, end_bit):
    L = W.lib()
    k, v = keys.clone(), vals.clone()
    k_alt, v_alt = torch.empty_like(k), torch.empty_like(v)
    status = torch.full((1,), 7, dtype=torch.int32, device=k.device)
    sel = ctypes.c_int(-1)
    L.check(L.ubs_radix_sort_pairs_noalloc(k.numel(), W.ptr(k), W.ptr(k_alt), W.ptr(v), W.ptr(v_alt), begin_bit,
                                           end_bit, ctypes.byref(sel), W.ptr(status), W.stream_of(k)),
            "radix_sort_pairs_noalloc")
    torch.cuda.synchronize()
    assert sel.value in (0, 1)
    return (k_alt, v_alt) if sel.value else (k, v), int(status.item())


@pytest.mark.parametrize("first,second", [((3_000_001, 0, 46), (100_003, 32, 45))])
def test_noalloc_sort_matches_workspace_sort(device, first, second):
    for n, b, e in (first, second):
        print(f"noalloc sort n={n} bits=[{b},{e})", flush=True)
        keys_h, vals_h = _keys(n, e, seed=n)
        keys, vals = torch.tensor(keys_h, device=device), torch.tensor(vals_h, device=device)
        (k1, v1), status = _noalloc(keys, vals, b, e)
        assert status == 0
        k2, v2 = W.radix_sort_pairs(keys.clone(), vals.clone(), b, e, check_status=True)
        assert torch.equal(k1, k2) and torch.equal(v1, v2), (n, b, e)
        mask = ((1 << e) - 1) ^ ((1 << b) - 1) if e < 64 else ~((1 << b) - 1)
        order = np.argsort(keys_h & mask, kind="stable")
        np.testing.assert_array_equal(k1.cpu().numpy(), keys_h[order])
        np.testing.assert_array_equal(v1.cpu().numpy(), vals_h[order])
