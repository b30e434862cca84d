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


def test_l1_grad_many_matches_single():
    """40 images of uneven sizes (two launches of at most 32 items), sizes not
    multiples of four: each gradient bit-identical to the one-image call."""
    if not gpu_available():
        pytest.skip("no GPU")
    import torch
    from differentiable_gaussian_mixture_renderer_amd import losses
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(11)
    rng = np.random.default_rng(5)
    shapes = [(3, int(h), int(w)) for h, w in rng.integers(1, 700, size=(40, 2))] + [(3, 1080, 1920)]
    imgs = [torch.rand(s, generator=g, device=dev) for s in shapes]
    tgts = [torch.rand(s, generator=g, device=dev) for s in shapes]
    tgts[0].view(-1)[::5] = imgs[0].view(-1)[::5]
    got = losses.l1_grad_many(imgs, tgts)
    for a, b, gg in zip(imgs, tgts, got):
        assert torch.equal(gg, losses.l1_grad(a, b))
