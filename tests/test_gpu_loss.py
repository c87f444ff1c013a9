"""GPU parity of the image loss kernels (ubs_image_loss_fwd/bwd through
gsplat.loss) against the fp64 oracle restating utils/loss_utils.py:45-80
(which the oracle's own tests pin: S(x, x) = 1, symmetry, the constant-image
closed form, finite differences).

Tolerances: loss relative 2e-5 (fp32 separable window vs the fp64 2-D
window); gradient relative L2 2e-4 (fp32 sums of 121 products per map).
"""
import numpy as np
import pytest
import torch

from gsplat import loss as Lmod
from oracle import ubs_oracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("shape", [(1, 40, 56, 3), (2, 33, 17, 3), (1, 16, 16, 1), (3, 7, 70, 2)])
@pytest.mark.parametrize("lam", [0.0, 0.2, 1.0])
def test_image_loss_matches_oracle(device, shape, lam):
    rng = np.random.Generator(np.random.PCG64(sum(shape) + int(lam * 10)))
    x = rng.uniform(0, 1, shape).astype(np.float32)
    y = np.clip(x + rng.normal(0, 0.2, shape), 0, 1).astype(np.float32)  # correlated, as renders and targets
    x[0, 0, 0, 0] = y[0, 0, 0, 0]  # one tie: sign(0) = 0
    want_loss, want_grad = O.image_loss(x.astype(np.float64), y.astype(np.float64), lam)
    tx = torch.tensor(x, device=device, requires_grad=True)
    got = Lmod.image_loss(tx, torch.tensor(y, device=device), lam)
    got.backward()
    assert abs(float(got) - want_loss) <= 2e-5 * abs(want_loss) + 1e-7, (float(got), want_loss)
    g = tx.grad.cpu().numpy()
    assert rel_l2(g, want_grad) < 2e-4, rel_l2(g, want_grad)
    assert g[0, 0, 0, 0] == pytest.approx(float(want_grad[0, 0, 0, 0]), abs=1e-6 * np.abs(want_grad).max())


def test_l1_loss_exact_cases(device):
    x = torch.tensor([[[0.0, 1.0], [2.0, 3.0]]], device=device).reshape(1, 2, 2, 1)
    y = torch.tensor([[[0.0, 3.0], [1.0, 3.0]]], device=device).reshape(1, 2, 2, 1)
    x.requires_grad_(True)
    l = Lmod.l1_loss(x, y)
    l.backward()
    assert float(l) == pytest.approx(0.75)
    assert x.grad.flatten().tolist() == [0.0, -0.25, 0.25, 0.0]
    # the upstream gradient scales the result; 3-D [H, W, C] input works too
    z = torch.rand(8, 9, 3, device=device, requires_grad=True)
    t = torch.rand(8, 9, 3, device=device)
    (3.0 * Lmod.image_loss(z, t, 0.2)).backward()
    g3 = z.grad.clone()
    z.grad = None
    Lmod.image_loss(z, t, 0.2).backward()
    torch.testing.assert_close(g3, 3.0 * z.grad)


def test_image_loss_rejects_bad_lambda(device):
    with pytest.raises(ValueError):
        Lmod.image_loss(torch.zeros(1, 4, 4, 3, device=device), torch.zeros(1, 4, 4, 3, device=device), 1.5)
