"""Fused L1 loss against torch: loss within float summation-order rounding,
gradient bit-identical to torch's sign(image - target) / n, including ties
(sign 0) and a size that is not a multiple of four."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(3, 37, 53), (3, 1080, 1920)])
def test_l1_loss_matches_torch(shape):
    if not gpu_available():
        pytest.skip("no GPU")
    import torch
    from differentiable_gaussian_mixture_renderer_amd import losses
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(7)
    img = torch.rand(shape, generator=g, device=dev)
    tgt = torch.rand(shape, generator=g, device=dev)
    tgt.view(-1)[::97] = img.view(-1)[::97]  # ties: gradient 0
    a = img.clone().requires_grad_(True)
    ref = (a - tgt).abs().mean()
    ref.backward()
    b = img.clone().requires_grad_(True)
    got = losses.l1_loss(b, tgt)
    got.backward()
    torch.cuda.synchronize()
    assert abs(float(got) - float(ref)) <= 1e-6 * float(ref)
    assert torch.equal(b.grad, a.grad)
    assert torch.equal(losses.l1_grad(img, tgt), a.grad)
