"""Optimiser tail of the training step, restated in torch (dtype-generic).

Restates
  * ``torch.nan_to_num`` + ``clip_grad_norm_(max_norm=1.0)``  train_2d_gaussian_rgb.py:205-209
    (total norm = sqrt(sum g^2) over all parameters, coefficient
    min(1, 1 / (norm + 1e-6)))
  * ``Adan.step`` / ``_single_tensor_adan``  utils/optimizer.py:105-262 with
    betas (0.98, 0.92, 0.99), eps 1e-8, weight decay 0, max_grad_norm 0
    (so its own clip factor is 1), no_prox False, per-group step counters that
    advance even for groups without gradients (:145-149).
  * ``setup_optimizer`` groups and learning rates  train_2d_gaussian_rgb.py:73-86
  * ``MeansCosineScheduler``                  train_2d_gaussian_rgb.py:58-70
    (closed form eta_min + (lr0 - eta_min) (1 + cos(pi t / T)) / 2 of
    CosineAnnealingLR, eta_min 1e-3).

Pinned by (tests/test_oracle_adan.py): the first step is sign descent
p - lr g / (|g| + eps) (m = (1-b1) g, n = (1-b3) g^2, diff = 0); the second
step against a scalar hand expansion of the published recurrences; the clip
against its definition; the schedule's end points.
"""
import math

import torch

PARAM_NAMES = ("means_2d", "pre_act_scales", "pre_act_rotation", "colours", "pre_act_phase",
               "pre_act_opacities", "curvature")
# (group name, member parameters, learning rate) -- train:75-82
GROUPS = (("means", ("means_2d",), None),
          ("scales", ("pre_act_scales",), 0.005),
          ("amplitude_phase", ("colours", "pre_act_phase"), 0.0025),
          ("opacity", ("pre_act_opacities",), 0.025),
          ("rotation", ("pre_act_rotation",), 0.001),
          ("curvature", ("curvature",), 0.001))
BETAS = (0.98, 0.92, 0.99)
EPS = 1e-8


def means_lr(t, num_itrs, lr0=0.01, eta_min=1e-3):
    T = max(1, num_itrs)
    return eta_min + (lr0 - eta_min) * (1 + math.cos(math.pi * t / T)) / 2


def clip_grads(grads, max_norm=1.0):
    """nan_to_num then global-norm clip; grads: dict name -> tensor or None."""
    out = {}
    for n, g in grads.items():
        out[n] = None if g is None else torch.nan_to_num(g, nan=0.0, posinf=0.0, neginf=0.0)
    gs = [g for g in out.values() if g is not None]
    total = torch.sqrt(sum((g.double() ** 2).sum() for g in gs)) if gs else torch.tensor(0.0)
    coef = min(1.0, max_norm / (float(total) + 1e-6))
    return {n: (None if g is None else g * coef) for n, g in out.items()}, float(total)


class AdanState:
    def __init__(self, lr0=0.01, num_itrs=2001):
        self.lr0 = lr0
        self.num_itrs = num_itrs
        self.step_count = 0          # shared by all groups (they advance together)
        self.sched_t = 0
        self.state = {}

    def group_lr(self, gname, lr):
        return means_lr(self.sched_t, self.num_itrs, self.lr0) if gname == "means" else lr

    def step(self, params, grads):
        """params, grads: dict name -> tensor (grads may be None).  Updates params in
        place (no autograd) and advances the schedule."""
        b1, b2, b3 = BETAS
        self.step_count += 1
        t = self.step_count
        bc1, bc2, bc3 = 1 - b1 ** t, 1 - b2 ** t, 1 - b3 ** t
        with torch.no_grad():
            for gname, members, lr in GROUPS:
                lr = self.group_lr(gname, lr)
                for n in members:
                    g = grads.get(n)
                    if g is None:
                        continue
                    p = params[n]
                    st = self.state.get(n)
                    if st is None:
                        st = dict(m=torch.zeros_like(p), v=torch.zeros_like(p),
                                  d=torch.zeros_like(p))
                        self.state[n] = st
                    if "npg" not in st or t == 1:
                        st["npg"] = -g.clone()
                    npg = st["npg"] + g                               # g_t - g_{t-1}
                    st["m"] = st["m"] * b1 + (1 - b1) * g
                    st["d"] = st["d"] * b2 + (1 - b2) * npg
                    upd = npg * b2 + g
                    st["v"] = st["v"] * b3 + (1 - b3) * upd * upd
                    denom = st["v"].sqrt() / math.sqrt(bc3) + EPS
                    p.sub_(lr / bc1 * st["m"] / denom)
                    p.sub_(lr * b2 / bc2 * st["d"] / denom)
                    st["npg"] = -g.clone()
        self.sched_t += 1
