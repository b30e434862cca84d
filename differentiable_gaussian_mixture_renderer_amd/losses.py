"""Photometric L1 loss of the training step as one fused pass each way
(dgmr_l1_loss_forward / dgmr_l1_loss_backward); the reference leaves the loss
to its caller.  `l1_loss(image, target)` equals `(image - target).abs().mean()`
and its gradient `sign(image - target) / n` (torch.sign semantics);
`l1_grad_many` is that gradient for many images in one launch
(dgmr_l1_loss_backward_many), for the views of a batch."""
from __future__ import annotations

import ctypes

import torch

from . import _common as C
from . import _native as N


class L1Loss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, target):
        image = image.contiguous()
        target = target.to(image.dtype).contiguous()
        if image.dtype != torch.float32 or image.shape != target.shape:
            raise ValueError("l1_loss: float32 image and target of one shape")
        lib = N.load()
        n = image.numel()
        count = ctypes.c_int32(0)
        N.check(lib.dgmr_l1_partials_size(n, ctypes.byref(count)), "dgmr_l1_partials_size")
        partials = torch.empty((count.value,), dtype=torch.float32, device=image.device)
        stream = C.stream_handle(image.device)
        N.check(lib.dgmr_l1_loss_forward(image.data_ptr(), target.data_ptr(), n, 1.0 / n, partials.data_ptr(), stream),
                "dgmr_l1_loss_forward")
        ctx.save_for_backward(image, target)
        return partials.sum()

    @staticmethod
    def backward(ctx, grad_output):
        image, target = ctx.saved_tensors
        n = image.numel()
        g = grad_output.reshape(1).to(torch.float32).contiguous()
        grad = torch.empty_like(image)
        N.check(N.load().dgmr_l1_loss_backward(image.data_ptr(), target.data_ptr(), n, 1.0 / n, g.data_ptr(), grad.data_ptr(),
                                               C.stream_handle(image.device)), "dgmr_l1_loss_backward")
        return grad, None


def l1_loss(image: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    return L1Loss.apply(image, target)


_ONE = {}


def l1_grad(image: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """dL/dimage of L = mean |image - target| in one pass (no loss value)."""
    image = image.contiguous()
    target = target.to(image.dtype).contiguous()
    one = _ONE.get(image.device)
    if one is None:
        one = _ONE[image.device] = torch.ones((1,), dtype=torch.float32, device=image.device)
    grad = torch.empty_like(image)
    n = image.numel()
    N.check(N.load().dgmr_l1_loss_backward(image.data_ptr(), target.data_ptr(), n, 1.0 / n, one.data_ptr(), grad.data_ptr(),
                                           C.stream_handle(image.device)), "dgmr_l1_loss_backward")
    return grad


def l1_items(images, targets, grads):
    """Descriptors (dgmr_l1_item) for dgmr_l1_loss_backward_many: item k
    writes grads[k] = sign(images[k] - targets[k]) / numel.  All float32,
    contiguous, on one device, 16-byte aligned."""
    if not images:
        return (N.L1Item * 0)()
    dev = images[0].device
    for i, t, g in zip(images, targets, grads):
        if (i.dtype != torch.float32 or t.dtype != torch.float32 or g.dtype != torch.float32 or i.shape != t.shape or i.shape != g.shape
                or not (i.is_contiguous() and t.is_contiguous() and g.is_contiguous())
                or i.device != dev or t.device != dev or g.device != dev):
            raise ValueError("l1_items: contiguous float32 images, targets and gradients of matching shapes on one device")
    one = _ONE.get(dev)
    if one is None:
        one = _ONE[dev] = torch.ones((1,), dtype=torch.float32, device=dev)
    items = (N.L1Item * len(images))()
    for k, (i, t, g) in enumerate(zip(images, targets, grads)):
        items[k].image, items[k].target, items[k].grad = i.data_ptr(), t.data_ptr(), g.data_ptr()
        items[k].grad_loss = one.data_ptr()
        items[k].n = i.numel()
        items[k].scale = 1.0 / max(i.numel(), 1)
    return items


def l1_grad_many(images, targets):
    """[l1_grad(i, t) for i, t in zip(images, targets)] in one launch."""
    images = [i.contiguous() for i in images]
    targets = [t.to(i.dtype).contiguous() for i, t in zip(images, targets)]
    if not images:
        return []
    grads = [torch.empty_like(i) for i in images]
    items = l1_items(images, targets, grads)
    N.check(N.load().dgmr_l1_loss_backward_many(items, len(images), C.stream_handle(images[0].device)), "dgmr_l1_loss_backward_many")
    return grads
