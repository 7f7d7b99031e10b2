"""CPU restatement of the learned warm-start network (test infrastructure only).

The reference has no learned part (its initialisation is random,
``model_2d_gaussian_rgb.py:40-80``); this module restates the network the HIP kernels of
``csrc/nn.hip`` implement, so they can be checked on the same inputs:

* activations NHWC, rounded to bfloat16 (round to nearest even) exactly where the GPU
  stores them: the tile rows of the input, every convolution output (after ReLU), the
  skip sums, the pooled maps; the head output stays FP32;
* convolutions are cross-correlations with "same" zero padding, filter rows
  ``w[n][tap * cin + c]`` with ``tap = ky * k + kx`` at offset ``(ky - k//2, kx - k//2)``;
* 2x2 average pooling to ``ceil(h/2) x ceil(w/2)`` (mean over the covered pixels) and
  nearest-neighbour upsampling back to the skip's size.

Gradients come from torch autograd in FP64 through the same graph, with the bfloat16
roundings of the forward treated as identities (straight-through), which is what the
GPU's adjoint computes up to its own bfloat16 rounding of the pre-activation
gradients.  How it is pinned: ``conv`` against a brute-force loop (``conv_bruteforce``),
pooling and upsampling against loops, and the composed gradient against central finite
differences of the unrounded network in FP64 (``tests/test_oracle_nn.py``).
"""
import math

import numpy as np
import torch
import torch.nn.functional as F

# (name, cin, cout, ksize, relu): the encoder-decoder, level 0 = one row per 16x16 tile
CHANNELS = (32, 64, 128, 256, 256)
SLOTS, SLOT_DIM = 8, 18
OUT_DIM = SLOTS * SLOT_DIM


def layers(in_ch=4):
    c = CHANNELS
    L = [("stem", in_ch * 256, c[0], 1, True), ("e0", c[0], c[0], 3, True)]
    for l in range(1, 4):
        L += [(f"e{l}a", c[l - 1], c[l], 3, True), (f"e{l}b", c[l], c[l], 3, True)]
    L += [("e4", c[3], c[4], 3, True)]
    for l in range(3, -1, -1):
        L += [(f"d{l}a", c[l + 1], c[l], 3, True), (f"d{l}b", c[l], c[l], 3, True)]
    L += [("head", c[0], OUT_DIM, 1, False)]
    return L


def bf16(t):
    return t.to(torch.bfloat16).to(t.dtype)


def tiles_from_image(img):
    """[B, C, H, W] -> [B, Ty, Tx, C*256], element (c*16 + dy)*16 + dx, zero outside."""
    B, C, H, W = img.shape
    Ty, Tx = -(-H // 16), -(-W // 16)
    pad = torch.zeros(B, C, Ty * 16, Tx * 16, dtype=img.dtype)
    pad[:, :, :H, :W] = img
    return pad.view(B, C, Ty, 16, Tx, 16).permute(0, 2, 4, 1, 3, 5).reshape(B, Ty, Tx, C * 256)


def conv(x, w, b, k):
    """x [B, H, W, cin], w [cout, k*k*cin], b [cout] or None -> [B, H, W, cout]."""
    cout, cin = w.shape[0], x.shape[-1]
    wk = w.reshape(cout, k, k, cin).permute(0, 3, 1, 2)
    y = F.conv2d(x.permute(0, 3, 1, 2), wk, b, padding=k // 2)
    return y.permute(0, 2, 3, 1)


def conv_bruteforce(x, w, b, k):
    """Direct loop over taps in FP64 numpy (pins ``conv``)."""
    x = np.asarray(x, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    B, H, W, cin = x.shape
    cout = w.shape[0]
    y = np.zeros((B, H, W, cout))
    for tap in range(k * k):
        dy, dx = tap // k - k // 2, tap % k - k // 2
        xs = np.zeros_like(x)
        ys0, ys1 = max(0, -dy), min(H, H - dy)
        xs0, xs1 = max(0, -dx), min(W, W - dx)
        xs[:, ys0:ys1, xs0:xs1] = x[:, ys0 + dy:ys1 + dy, xs0 + dx:xs1 + dx]
        y += xs @ w[:, tap * cin:(tap + 1) * cin].T
    if b is not None:
        y += np.asarray(b, dtype=np.float64)
    return y


def pool2(x):
    B, H, W, C = x.shape
    Ho, Wo = -(-H // 2), -(-W // 2)
    pad = torch.zeros(B, 2 * Ho, 2 * Wo, C, dtype=x.dtype)
    cnt = torch.zeros(1, 2 * Ho, 2 * Wo, 1, dtype=x.dtype)
    pad[:, :H, :W] = x
    cnt[:, :H, :W] = 1
    s = pad.view(B, Ho, 2, Wo, 2, C).sum((2, 4))
    n = cnt.view(1, Ho, 2, Wo, 2, 1).sum((2, 4))
    return s / n


def up2(x, H, W):
    return x.repeat_interleave(2, 1).repeat_interleave(2, 2)[:, :H, :W]


class _Round(torch.autograd.Function):
    """bfloat16 rounding in the forward, identity in the backward."""

    @staticmethod
    def forward(ctx, t):
        return bf16(t)

    @staticmethod
    def backward(ctx, g):
        return g


def forward(img, params, rounding=True):
    """img [B, C, H, W]; params {name: (w [cout, k*k*cin], b [cout])} -> [B, Ty, Tx, 144]."""
    R = _Round.apply if rounding else (lambda t: t)
    spec = {n: (ci, co, k, a) for n, ci, co, k, a in layers(img.shape[1])}

    def cv(name, x):
        _, _, k, act = spec[name]
        w, b = params[name]
        y = conv(x, R(w), b, k)
        return torch.relu(y) if act else y

    x0 = R(tiles_from_image(img))
    s = R(cv("stem", x0))
    e = [R(cv("e0", s))]
    h = e[0]
    for l in range(1, 4):
        h = R(pool2(h))
        h = R(cv(f"e{l}a", h))
        h = R(cv(f"e{l}b", h))
        e.append(h)
    h = R(pool2(h))
    h = R(cv("e4", h))
    for l in range(3, -1, -1):
        H, W = e[l].shape[1:3]
        u = up2(h, H, W)
        t = R(R(cv(f"d{l}a", u)) + e[l])
        h = R(cv(f"d{l}b", t))
    return cv("head", h)


def init_params(in_ch=4, seed=0, dtype=torch.float32):
    """He-normal filters, zero biases, head scaled by 0.1 (seeded)."""
    g = torch.Generator().manual_seed(seed)
    out = {}
    for name, ci, co, k, act in layers(in_ch):
        fan_in = ci * k * k
        w = torch.randn(co, k * k * ci, generator=g, dtype=torch.float64) * math.sqrt(2.0 / fan_in)
        if name == "head":
            w = w * 0.1
        out[name] = (w.to(dtype), torch.zeros(co, dtype=dtype))
    return out


def loss_and_grads(img, params, target, weight, scale, rounding=True):
    """FP64 weighted squared error of the network output and its parameter gradients."""
    p64 = {n: (w.double().clone().requires_grad_(), b.double().clone().requires_grad_())
           for n, (w, b) in params.items()}
    out = forward(img.double(), p64, rounding=rounding)
    L = (scale * weight.double() * (out - target.double()) ** 2).sum()
    L.backward()
    return float(L.detach()), out.detach(), {n: (w.grad, b.grad) for n, (w, b) in p64.items()}
