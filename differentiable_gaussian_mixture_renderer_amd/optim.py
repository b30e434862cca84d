"""Fused Adam over the Gaussian parameter tensors (dgmr_adam_step).

`FusedAdam(params, lr=..., betas=(0.9, 0.999), eps=1e-8, weight_decay=0)` is a
torch.optim.Optimizer with torch.optim.Adam's update rule (L2 weight decay,
bias-corrected moments).  Each step is one kernel launch per float32 CUDA
tensor, reading parameter, gradient and both moments once and writing
parameter and moments once.  The reference has no optimiser of its own.
"""
from __future__ import annotations

import ctypes

import torch

from . import _common as C
from . import _native as N


class FusedAdam(torch.optim.Optimizer):
    def __init__(self, params, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8, weight_decay: float = 0.0):
        if lr < 0 or eps <= 0 or not (0 <= betas[0] < 1 and 0 <= betas[1] < 1):
            raise ValueError("invalid Adam hyper-parameters")
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps, weight_decay=weight_decay))

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        lib = N.load()
        for group in self.param_groups:
            table, keep, steps = [], [], set()
            for p in group["params"]:
                if p.grad is None:
                    continue
                if p.dtype != torch.float32 or not p.is_cuda or not p.is_contiguous():
                    raise TypeError("FusedAdam takes contiguous float32 CUDA parameters")
                st = self.state[p]
                if not st:
                    st["step"] = 0
                    st["exp_avg"] = torch.zeros_like(p)
                    st["exp_avg_sq"] = torch.zeros_like(p)
                st["step"] += 1
                steps.add(st["step"])
                g = p.grad.contiguous()
                keep.append(g)
                table.append(N.AdamTensor(p.data_ptr(), g.data_ptr(), st["exp_avg"].data_ptr(), st["exp_avg_sq"].data_ptr(), p.numel(),
                                          float(group["lr"])))
            if not table:
                continue
            if len(steps) != 1:
                raise RuntimeError("FusedAdam: parameters of one group must share a step count")
            arr = (N.AdamTensor * len(table))(*table)
            dev = group["params"][0].device
            b1, b2 = group["betas"]
            N.check(lib.dgmr_adam_step(arr, len(table), float(b1), float(b2), float(group["eps"]), float(group["weight_decay"]),
                                       int(steps.pop()), C.stream_handle(dev)), "dgmr_adam_step")
        return loss
