"""Adan optimiser with the reference's interface (utils/optimizer.py:25-204).

``Adan(params, lr, betas, eps, weight_decay, max_grad_norm, no_prox, foreach)``
follows the published recurrences (Xie et al. 2022) exactly as the reference's
``_single_tensor_adan`` (utils/optimizer.py:207-262) applies them:

    g      <- g * clip
    diff   <- g - g_prev
    m      <- b1 m + (1 - b1) g
    d      <- b2 d + (1 - b2) diff
    n      <- b3 n + (1 - b3) (g + b2 diff)^2
    p      <- p - lr / (1 - b1^t) * m / (sqrt(n) / sqrt(1 - b3^t) + eps)
             - lr * b2 / (1 - b2^t) * d / (same denominator)
    p      <- p / (1 + lr * wd)            (no_prox False)

with one step counter per group that advances on every call, and g_prev
initialised to g on the first step a parameter has a gradient.  This PyTorch
version serves the reference-interface path and the PyTorch baseline; the fused
training step runs the same update in one gfx950 kernel.
"""
import math

import torch
from torch.optim.optimizer import Optimizer


class Adan(Optimizer):
    def __init__(self, params, lr=1e-3, betas=(0.98, 0.92, 0.99), eps=1e-8, weight_decay=0.0,
                 max_grad_norm=0.0, no_prox=False, foreach=True):
        if max_grad_norm < 0 or lr < 0 or eps < 0:
            raise ValueError("invalid Adan hyper-parameter")
        for b in betas:
            if not 0.0 <= b < 1.0:
                raise ValueError(f"Invalid beta parameter: {b}")
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps, weight_decay=weight_decay,
                                      max_grad_norm=max_grad_norm, no_prox=no_prox,
                                      foreach=foreach))

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        clip = 1.0
        if self.defaults["max_grad_norm"] > 0:
            norms = [p.grad.pow(2).sum() for g in self.param_groups for p in g["params"]
                     if p.grad is not None]
            total = torch.sqrt(torch.stack(norms).sum()) if norms else torch.zeros(())
            clip = float(torch.clamp(self.defaults["max_grad_norm"] /
                                     (total + self.param_groups[-1]["eps"]), max=1.0))
        for group in self.param_groups:
            b1, b2, b3 = group["betas"]
            group["step"] = group.get("step", 0) + 1
            t = group["step"]
            bc1, bc2, bc3 = 1 - b1 ** t, 1 - b2 ** t, 1 - b3 ** t
            lr, eps, wd = group["lr"], group["eps"], group["weight_decay"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                g = p.grad
                if clip != 1.0:
                    g.mul_(clip)
                st = self.state[p]
                if len(st) == 0:
                    st["exp_avg"] = torch.zeros_like(p)
                    st["exp_avg_sq"] = torch.zeros_like(p)
                    st["exp_avg_diff"] = torch.zeros_like(p)
                if "neg_pre_grad" not in st or t == 1:
                    st["neg_pre_grad"] = -g.clone()
                diff = st["neg_pre_grad"].add_(g)                  # g - g_prev
                st["exp_avg"].mul_(b1).add_(g, alpha=1 - b1)
                st["exp_avg_diff"].mul_(b2).add_(diff, alpha=1 - b2)
                upd = diff.mul_(b2).add_(g)
                st["exp_avg_sq"].mul_(b3).addcmul_(upd, upd, value=1 - b3)
                denom = (st["exp_avg_sq"].sqrt() / math.sqrt(bc3)).add_(eps)
                if group["no_prox"]:
                    p.mul_(1 - lr * wd)
                p.addcdiv_(st["exp_avg"], denom, value=-lr / bc1)
                p.addcdiv_(st["exp_avg_diff"], denom, value=-lr * b2 / bc2)
                if not group["no_prox"]:
                    p.div_(1 + lr * wd)
                st["neg_pre_grad"].copy_(g).neg_()
        return loss
