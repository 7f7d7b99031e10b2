"""Pins oracle.adan (utils/optimizer.py:105-262, train_2d_gaussian_rgb.py:58-86,205-211).

Each check is computed a second way, independently of the oracle's tensor code:
  * the first Adan step is sign descent p - lr g / (|g| + eps) (the gradient difference
    is zero at step 1 and the bias corrections cancel the (1 - beta) factors);
  * the second and third steps against a scalar, per-element hand expansion of the
    recurrences of _single_tensor_adan (utils/optimizer.py:236-262), written with
    Python floats;
  * nan_to_num + clip against torch.nn.utils.clip_grad_norm_ (torch's own routine);
  * the means schedule against torch.optim.lr_scheduler.CosineAnnealingLR driven the
    way MeansCosineScheduler drives it (train_2d_gaussian_rgb.py:58-70), including
    its end points lr0 and eta_min.
"""
import math

import numpy as np
import torch

from oracle import adan

B1, B2, B3 = adan.BETAS
EPS = adan.EPS


def _params(seed):
    rng = np.random.default_rng(seed)
    shapes = {"means_2d": (5, 2), "pre_act_scales": (5, 2), "pre_act_rotation": (5,),
              "colours": (5, 3), "pre_act_phase": (5, 3), "pre_act_opacities": (5,),
              "curvature": (5,)}
    p = {n: torch.as_tensor(rng.normal(size=s)) for n, s in shapes.items()}
    gs = [{n: torch.as_tensor(rng.normal(size=s)) for n, s in shapes.items()} for _ in range(3)]
    return p, gs


LR = {"means_2d": None, "pre_act_scales": 0.005, "pre_act_rotation": 0.001, "colours": 0.0025,
      "pre_act_phase": 0.0025, "pre_act_opacities": 0.025, "curvature": 0.001}


def test_first_step_is_sign_descent():
    p, gs = _params(0)
    st = adan.AdanState(lr0=0.01, num_itrs=100)
    q = {n: v.clone() for n, v in p.items()}
    st.step(q, gs[0])
    for n in p:
        lr = 0.01 if LR[n] is None else LR[n]
        expect = p[n] - lr * gs[0][n] / (gs[0][n].abs() + EPS)
        torch.testing.assert_close(q[n], expect, rtol=1e-12, atol=1e-15)


def _scalar_adan(p0, grads, lrs):
    """One element through utils/optimizer.py:236-262 with Python floats; lrs[t] is the
    learning rate at step t + 1."""
    p, m, v, d = float(p0), 0.0, 0.0, 0.0
    neg = None
    for t, g in enumerate(grads, start=1):
        g = float(g)
        if t == 1:
            neg = -g
        bc1, bc2, bc3 = 1 - B1 ** t, 1 - B2 ** t, 1 - B3 ** t
        diff = neg + g
        m = m * B1 + (1 - B1) * g
        d = d * B2 + (1 - B2) * diff
        u = diff * B2 + g
        v = v * B3 + (1 - B3) * u * u
        den = math.sqrt(v) / math.sqrt(bc3) + EPS
        p = p - lrs[t - 1] / bc1 * m / den
        p = p - lrs[t - 1] * B2 / bc2 * d / den
        neg = -g
    return p


def test_three_steps_match_scalar_expansion():
    p, gs = _params(1)
    T = 10
    st = adan.AdanState(lr0=0.01, num_itrs=T)
    q = {n: v.clone() for n, v in p.items()}
    for g in gs:
        st.step(q, g)
    means_lrs = [1e-3 + (0.01 - 1e-3) * (1 + math.cos(math.pi * t / T)) / 2 for t in range(3)]
    for n in p:
        lrs = means_lrs if LR[n] is None else [LR[n]] * 3
        flat0 = p[n].reshape(-1)
        got = q[n].reshape(-1)
        for i in range(flat0.numel()):
            exp = _scalar_adan(flat0[i], [g[n].reshape(-1)[i] for g in gs], lrs)
            assert abs(float(got[i]) - exp) <= 1e-12 * max(1.0, abs(exp)), (n, i)


def test_frozen_group_keeps_shared_step_count():
    """A parameter without a gradient (curvature during the warm-up) is skipped, while the
    step counter of every group advances (utils/optimizer.py:145-149): when it unfreezes at
    step 3, its first update uses bias corrections of step 3 and a zero difference."""
    p, gs = _params(2)
    st = adan.AdanState(lr0=0.01, num_itrs=100)
    q = {n: v.clone() for n, v in p.items()}
    for i, g in enumerate(gs):
        g = dict(g)
        if i < 2:
            g["curvature"] = None
        st.step(q, g)
    torch.testing.assert_close(q["curvature"],
                               p["curvature"] - _curv_update(p["curvature"], gs[2]["curvature"]),
                               rtol=1e-12, atol=1e-15)


def _curv_update(p, g, t=3, lr=0.001):
    bc1, bc2, bc3 = 1 - B1 ** t, 1 - B2 ** t, 1 - B3 ** t
    m = (1 - B1) * g
    v = (1 - B3) * g * g
    den = v.sqrt() / math.sqrt(bc3) + EPS
    return lr / bc1 * m / den


def test_clip_matches_torch_clip_grad_norm():
    rng = np.random.default_rng(3)
    raw = {f"p{i}": rng.normal(size=(7, 3)) * 0.4 for i in range(4)}
    raw["p0"][1, 2] = np.nan
    raw["p2"][0, 0] = np.inf
    grads = {n: torch.as_tensor(v) for n, v in raw.items()}
    got, total = adan.clip_grads(grads, max_norm=1.0)
    ps = []
    for n, v in raw.items():
        t = torch.zeros_like(torch.as_tensor(v), requires_grad=True)
        t.grad = torch.nan_to_num(torch.as_tensor(v), nan=0.0, posinf=0.0, neginf=0.0)
        ps.append(t)
    tn = torch.nn.utils.clip_grad_norm_(ps, max_norm=1.0)
    assert abs(total - float(tn)) <= 1e-12 * float(tn)
    assert float(tn) > 1.0            # the clip is active in this case
    for t, n in zip(ps, raw):
        torch.testing.assert_close(got[n], t.grad, rtol=1e-12, atol=1e-15)


def test_means_schedule_matches_torch_cosine_annealing():
    T = 37
    w = torch.zeros(1, requires_grad=True)
    opt = torch.optim.SGD([w], lr=0.01)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=T, eta_min=1e-3)
    lrs = [opt.param_groups[0]["lr"]]
    for _ in range(T):
        opt.step()
        sched.step()
        lrs.append(sched.get_last_lr()[0])
    for t, lr in enumerate(lrs):
        assert abs(adan.means_lr(t, T) - lr) <= 1e-12, t
    assert abs(adan.means_lr(0, T) - 0.01) <= 1e-15
    assert abs(adan.means_lr(T, T) - 1e-3) <= 1e-15
