"""Learned warm start for the hologram fit.

The reference starts every fit from random Gaussians (model_2d_gaussian_rgb.py:40-80)
and optimises them (train_2d_gaussian_rgb.py:187-211).  Here a small convolutional
encoder-decoder predicts, from the target image and depth map, 8 Gaussians for every
16x16 tile; those replace 8 x tiles of the random initial Gaussians, and the same fused
fit then runs from there.  The plain initialisation stays the fallback and the yardstick:
``time_to_psnr`` measures how long each start needs to reach the plain fit's final
per-plane PSNR.

Network (all on hand-written bf16 MFMA kernels with FP32 accumulation, ``csrc/nn.hip``):
one row per 16x16 tile (the tile's 4 x 256 pixels of RGB + depth), a 1x1 stem to 32
channels, a U-Net over the tile grid with 4 down stages (2x2 average pooling, 3x3
convolutions, 32-64-128-256-256 channels), 4 up stages (nearest upsampling, 3x3
convolutions, additive skips) and a 1x1 head to 8 slots x 18 values per tile;
3.56 M parameters.  Training, Adam on FP32 master weights with bf16 copies:
  1. supervised: weighted squared error against per-tile encodings of the Gaussians
     that plain fits produce, from a head whose bias is the training set's mean slot
     encoding (so the untrained network is the image-independent "mean start");
  2. render-loss fine-tuning (``finetune``): the decoded slots are rendered by the fused
     step and the fit's own loss is back-propagated through the renderer, the decode and
     the network, so the network learns starts that already render close to the
     multi-plane targets.
The fine-tuned network's start uses every slot (``all_slot_params``).  For a network
trained only on the supervised targets, ``warm_params`` decodes the slots it marks
present and takes a predicted phase only where its (cos, sin) vector is long enough to
be confident (|v| >= 1/2), the mean slot's phase elsewhere: phases of fitted Gaussians
are largely unpredictable from the image, and scattered phases slow the fit down.

Slot encoding (per Gaussian, tile origin (16 tx, 16 ty)):
  0-1   position in the tile, 2 (m - 16 t) / 16 - 1
  2-3   pre-activation log scales (scales = exp(.) + 0.1)
  4-5   cos 2 theta, sin 2 theta (rotation is pi-periodic)
  6-8   per channel amplitude |a|
  9-14  per channel cos(phi), sin(phi) of the phase (of a e^{i phi}, so a sign of a is
        folded into phi)
  15    tanh of the curvature parameter (curvature / its scale-aware bound)
  16    opacity
  17    1 if the slot holds a Gaussian
Slots hold a tile's Gaussians in decreasing opacity x amplitude^2 x area.  An empty
slot's target is zero amplitude, opacity and presence, its other values are free (zero
weight).  Decoded slots with presence below 1/2 keep the random initial Gaussian.
"""
import math

import numpy as np
import torch

from . import _lib

CHANNELS = (32, 64, 128, 256, 256)
SLOTS, SLOT_DIM = 8, 18
OUT_DIM = SLOTS * SLOT_DIM
IN_CH = 4  # RGB target + depth


def layer_table(in_ch=IN_CH):
    """(name, cin, cout, ksize, relu) in execution order of the forward."""
    c = CHANNELS
    t = [("stem", in_ch * 256, c[0], 1, True), ("e0", c[0], c[0], 3, True)]
    for l in (1, 2, 3):
        t.append((f"e{l}a", c[l - 1], c[l], 3, True))
        t.append((f"e{l}b", c[l], c[l], 3, True))
    t.append(("e4", c[3], c[4], 3, True))
    for l in (3, 2, 1, 0):
        t.append((f"d{l}a", c[l + 1], c[l], 3, True))
        t.append((f"d{l}b", c[l], c[l], 3, True))
    t.append(("head", c[0], OUT_DIM, 1, False))
    return t


def _levels(Ty, Tx):
    sizes = [(Ty, Tx)]
    for _ in range(4):
        h, w = sizes[-1]
        sizes.append(((h + 1) // 2, (w + 1) // 2))
    return sizes


class WarmStartNet:
    """Parameters, activations and the forward / backward / Adam sequence of the net."""

    def __init__(self, device="cuda", seed=0, in_ch=IN_CH):
        self.dev = torch.device(device)
        self.in_ch = in_ch
        self.layers = layer_table(in_ch)
        self.spec = {n: (ci, co, k, a) for n, ci, co, k, a in self.layers}
        off = 0
        self.offsets = {}
        for n, ci, co, k, a in self.layers:
            nw = co * k * k * ci
            self.offsets[n] = (off, off + nw)
            off += nw + co
            off = (off + 7) // 8 * 8          # keeps every filter 16-byte aligned
        self.numel = off
        self.n_params = sum(co * k * k * ci + co for _, ci, co, k, _ in self.layers)
        dev = self.dev
        self.w32 = torch.zeros(self.numel, device=dev)
        self.wb = torch.zeros(self.numel, dtype=torch.bfloat16, device=dev)
        self.grad = torch.zeros(self.numel, device=dev)
        self.m = torch.zeros(self.numel, device=dev)
        self.v = torch.zeros(self.numel, device=dev)
        self.flip = {n: torch.empty(co * k * k * ci, dtype=torch.bfloat16, device=dev)
                     for n, ci, co, k, _ in self.layers}
        self.steps = 0
        self._acts = {}
        self.loss = torch.zeros(1, device=dev)
        g = torch.Generator().manual_seed(seed)
        init = torch.zeros(self.numel)
        for n, ci, co, k, _ in self.layers:
            a, b = self.offsets[n]
            w = torch.randn(co * k * k * ci, generator=g) * math.sqrt(2.0 / (ci * k * k))
            init[a:b] = w * (0.1 if n == "head" else 1.0)
        self.w32.copy_(init.to(dev))
        self._sync_bf16()

    # ------------------------------------------------------------------ parameters
    def _sync_bf16(self):
        L = _lib.lib()
        _lib.check(L.cvqpg_nn_to_bf16(_lib.ptr(self.w32), self.numel, _lib.ptr(self.wb),
                                      _lib.stream_handle(self.dev)))
        self._refresh_flips()

    def _refresh_flips(self):
        L = _lib.lib()
        for n, ci, co, k, _ in self.layers:
            if n == "stem":
                continue  # its data gradient is never needed
            _lib.check(L.cvqpg_nn_flip_filter(_lib.ptr(self.wb_view(n)), co, ci, k,
                                              _lib.ptr(self.flip[n]), _lib.stream_handle(self.dev)))

    def view(self, buf, name):
        ci, co, k, _ = self.spec[name]
        a, b = self.offsets[name]
        return buf[a:b].view(co, k * k * ci), buf[b:b + co]

    def wb_view(self, name):
        a, b = self.offsets[name]
        return self.wb[a:b]

    def set_output_bias(self, vec):
        """Head bias = ``vec`` [144]: with the head's small initial filter the untrained
        network then predicts about ``vec`` everywhere and learns image-dependent
        deviations from it."""
        _, bv = self.view(self.w32, "head")
        bv.copy_(vec.to(bv))
        self._sync_bf16()

    def state_dict(self):
        return {n: tuple(t.detach().cpu().clone() for t in self.view(self.w32, n))
                for n, *_ in self.layers}

    def load_state_dict(self, params):
        for n, (w, b) in params.items():
            wv, bv = self.view(self.w32, n)
            wv.copy_(w)
            bv.copy_(b)
        self._sync_bf16()

    # ------------------------------------------------------------------ buffers
    def activations(self, B, Ty, Tx):
        key = (B, Ty, Tx)
        if key in self._acts:
            return self._acts[key]
        dev, c = self.dev, CHANNELS
        lv = _levels(Ty, Tx)
        bf = lambda h, w, ch: torch.empty(B, h, w, ch, dtype=torch.bfloat16, device=dev)
        f32 = lambda h, w, ch: torch.empty(B, h, w, ch, device=dev)
        A = {"B": B, "lv": lv}
        A["x0"] = bf(*lv[0], self.in_ch * 256)
        A["s"] = bf(*lv[0], c[0])
        A["e0"] = bf(*lv[0], c[0])
        for l in (1, 2, 3):
            A[f"p{l}"] = bf(*lv[l], c[l - 1])
            A[f"e{l}a"] = bf(*lv[l], c[l])
            A[f"e{l}b"] = bf(*lv[l], c[l])
        A["p4"] = bf(*lv[4], c[3])
        A["e4"] = bf(*lv[4], c[4])
        for l in (3, 2, 1, 0):
            A[f"u{l}"] = bf(*lv[l], c[l + 1])
            A[f"d{l}a"] = bf(*lv[l], c[l])
            A[f"t{l}"] = bf(*lv[l], c[l])
            A[f"d{l}b"] = bf(*lv[l], c[l])
        A["out"] = f32(*lv[0], OUT_DIM)
        # gradients (FP32); the gradient of a skip source aliases that of the skip sum
        G = {"out": f32(*lv[0], OUT_DIM), "d0b": f32(*lv[0], c[0])}
        for l in (0, 1, 2, 3):
            G[f"t{l}"] = f32(*lv[l], c[l])
            G[f"u{l}"] = f32(*lv[l], c[l + 1])
        for l in (1, 2, 3):
            G[f"d{l}b"] = f32(*lv[l], c[l])
            G[f"e{l}a"] = f32(*lv[l], c[l])
            G[f"p{l}"] = f32(*lv[l], c[l - 1])
        G["e4"] = f32(*lv[4], c[4])
        G["p4"] = f32(*lv[4], c[3])
        G["s"] = f32(*lv[0], c[0])
        G["e0"], G["e1b"], G["e2b"], G["e3b"] = G["t0"], G["t1"], G["t2"], G["t3"]
        A["grad"] = G
        big = max(B * h * w * ch for (h, w), ch in zip(lv, c))
        big = max(big, B * lv[0][0] * lv[0][1] * OUT_DIM)
        A["dpre"] = torch.empty(big, dtype=torch.bfloat16, device=dev)
        lvl = {"stem": 0, "e0": 0, "head": 0, "e4": 4}
        for l in (1, 2, 3):
            lvl[f"e{l}a"] = lvl[f"e{l}b"] = l
        for l in (0, 1, 2, 3):
            lvl[f"d{l}a"] = lvl[f"d{l}b"] = l
        L = _lib.lib()
        nws = max(int(L.cvqpg_nn_wgrad_workspace_size(B, *lv[lvl[n]], ci, co, k))
                  for n, ci, co, k, _ in self.layers)
        A["wgrad_ws"] = torch.empty(nws, dtype=torch.uint8, device=dev)
        n_out = B * lv[0][0] * lv[0][1] * OUT_DIM
        A["loss_ws"] = torch.empty(int(_lib.lib().cvqpg_nn_wmse_workspace_size(n_out)),
                                   dtype=torch.uint8, device=dev)
        self._acts[key] = A
        return A

    # ------------------------------------------------------------------ forward
    def _conv(self, name, x, B, h, w, yb=None, y2=None, addend=None, yf=None):
        ci, co, k, act = self.spec[name]
        wv = self.wb_view(name)
        _, bv = self.view(self.w32, name)
        _lib.check(_lib.lib().cvqpg_nn_conv(
            _lib.ptr(x), B, h, w, ci, _lib.ptr(wv), _lib.ptr(bv), co, k, 1 if act else 0,
            _lib.ptr(addend), _lib.ptr(yb), _lib.ptr(y2), _lib.ptr(yf),
            _lib.stream_handle(self.dev)))

    def forward(self, img):
        """img [B, 4, H, W] FP32 (RGB target in [0, 1] and depth in [0, 1]) ->
        [B, ceil(H/16), ceil(W/16), 144] FP32 slot encodings.  No host synchronisation."""
        img = img.contiguous().float()
        B, C, H, W = img.shape
        assert C == self.in_ch
        Ty, Tx = -(-H // 16), -(-W // 16)
        A = self.activations(B, Ty, Tx)
        lv = A["lv"]
        L, s = _lib.lib(), _lib.stream_handle(self.dev)
        _lib.check(L.cvqpg_nn_tiles_from_image(_lib.ptr(img), B, C, H, W, _lib.ptr(A["x0"]), s))
        h0, w0 = lv[0]
        self._conv("stem", A["x0"], B, h0, w0, yb=A["s"])
        self._conv("e0", A["s"], B, h0, w0, yb=A["e0"])
        prev = A["e0"]
        for l in (1, 2, 3):
            (hp, wp), (h, w) = lv[l - 1], lv[l]
            _lib.check(L.cvqpg_nn_pool2(_lib.ptr(prev), B, hp, wp, CHANNELS[l - 1],
                                        _lib.ptr(A[f"p{l}"]), s))
            self._conv(f"e{l}a", A[f"p{l}"], B, h, w, yb=A[f"e{l}a"])
            self._conv(f"e{l}b", A[f"e{l}a"], B, h, w, yb=A[f"e{l}b"])
            prev = A[f"e{l}b"]
        (h3, w3), (h4, w4) = lv[3], lv[4]
        _lib.check(L.cvqpg_nn_pool2(_lib.ptr(prev), B, h3, w3, CHANNELS[3], _lib.ptr(A["p4"]), s))
        self._conv("e4", A["p4"], B, h4, w4, yb=A["e4"])
        prev = A["e4"]
        for l in (3, 2, 1, 0):
            (hs, ws), (h, w) = lv[l + 1], lv[l]
            _lib.check(L.cvqpg_nn_upsample2(_lib.ptr(prev), B, hs, ws, CHANNELS[l + 1], h, w,
                                            _lib.ptr(A[f"u{l}"]), s))
            skip = A["e0"] if l == 0 else A[f"e{l}b"]
            self._conv(f"d{l}a", A[f"u{l}"], B, h, w, yb=A[f"d{l}a"], y2=A[f"t{l}"], addend=skip)
            self._conv(f"d{l}b", A[f"t{l}"], B, h, w, yb=A[f"d{l}b"])
            prev = A[f"d{l}b"]
        self._conv("head", prev, B, h0, w0, yf=A["out"])
        return A["out"]

    # ------------------------------------------------------------------ backward
    def _conv_bwd(self, name, x, post, dpost, A, h, w, dx=None, accumulate=False):
        """Weight and bias gradients of one convolution and (optionally) its data gradient."""
        ci, co, k, act = self.spec[name]
        B = A["B"]
        L, s = _lib.lib(), _lib.stream_handle(self.dev)
        n = B * h * w * co
        dpre = A["dpre"][:n]
        _lib.check(L.cvqpg_nn_relu_backward(_lib.ptr(dpost), _lib.ptr(post) if act else None, n,
                                            _lib.ptr(dpre), s))
        gw, gb = self.view(self.grad, name)
        _lib.check(L.cvqpg_nn_conv_wgrad(_lib.ptr(x), B, h, w, ci, _lib.ptr(dpre), co, k,
                                         _lib.ptr(gw), _lib.ptr(gb), _lib.ptr(A["wgrad_ws"]),
                                         A["wgrad_ws"].numel(), s))
        if dx is not None:
            _lib.check(L.cvqpg_nn_conv(_lib.ptr(dpre), B, h, w, co, _lib.ptr(self.flip[name]), None,
                                       ci, k, 2 if accumulate else 0, None, None, None,
                                       _lib.ptr(dx), s))

    def backward(self, A):
        """Parameter gradients into ``self.grad`` from the output gradient in A['grad']['out']
        (after ``forward`` on the same buffers)."""
        G, lv, B = A["grad"], A["lv"], A["B"]
        L, s = _lib.lib(), _lib.stream_handle(self.dev)
        h0, w0 = lv[0]
        self._conv_bwd("head", A["d0b"], None, G["out"], A, h0, w0, dx=G["d0b"])
        dcur = G["d0b"]
        for l in (0, 1, 2, 3):
            h, w = lv[l]
            # d{l}b: t{l} -> d{l}b ; the skip sum t = d{l}a + skip passes its gradient to both
            self._conv_bwd(f"d{l}b", A[f"t{l}"], A[f"d{l}b"], dcur, A, h, w, dx=G[f"t{l}"])
            self._conv_bwd(f"d{l}a", A[f"u{l}"], A[f"d{l}a"], G[f"t{l}"], A, h, w, dx=G[f"u{l}"])
            hs, ws = lv[l + 1]
            nxt = G["e4"] if l == 3 else G[f"d{l + 1}b"]
            _lib.check(L.cvqpg_nn_upsample2_backward(_lib.ptr(G[f"u{l}"]), B, hs, ws,
                                                     CHANNELS[l + 1], h, w, _lib.ptr(nxt), 0, s))
            dcur = nxt
        h4, w4 = lv[4]
        self._conv_bwd("e4", A["p4"], A["e4"], G["e4"], A, h4, w4, dx=G["p4"])
        _lib.check(L.cvqpg_nn_pool2_backward(_lib.ptr(G["p4"]), B, *lv[3], CHANNELS[3],
                                             _lib.ptr(G["e3b"]), 1, s))
        for l in (3, 2, 1):
            h, w = lv[l]
            self._conv_bwd(f"e{l}b", A[f"e{l}a"], A[f"e{l}b"], G[f"e{l}b"], A, h, w,
                           dx=G[f"e{l}a"])
            self._conv_bwd(f"e{l}a", A[f"p{l}"], A[f"e{l}a"], G[f"e{l}a"], A, h, w, dx=G[f"p{l}"])
            below = G["e0"] if l == 1 else G[f"e{l - 1}b"]
            _lib.check(L.cvqpg_nn_pool2_backward(_lib.ptr(G[f"p{l}"]), B, *lv[l - 1],
                                                 CHANNELS[l - 1], _lib.ptr(below), 1, s))
        self._conv_bwd("e0", A["s"], A["e0"], G["e0"], A, h0, w0, dx=G["s"])
        self._conv_bwd("stem", A["x0"], A["s"], G["s"], A, h0, w0)

    def loss_backward(self, A, target, weight, scale):
        """Weighted squared error of A['out'] (device scalar in self.loss) and its gradient."""
        n = A["out"].numel()
        _lib.check(_lib.lib().cvqpg_nn_wmse(
            _lib.ptr(A["out"]), _lib.ptr(target), _lib.ptr(weight), n, float(scale),
            _lib.ptr(A["grad"]["out"]), _lib.ptr(A["loss_ws"]), A["loss_ws"].numel(),
            _lib.ptr(self.loss), _lib.stream_handle(self.dev)))

    def adam(self, lr, betas=(0.9, 0.999), eps=1e-8):
        self.steps += 1
        _lib.check(_lib.lib().cvqpg_nn_adam(
            _lib.ptr(self.w32), _lib.ptr(self.wb), _lib.ptr(self.grad), _lib.ptr(self.m),
            _lib.ptr(self.v), self.numel, float(lr), betas[0], betas[1], eps, self.steps,
            _lib.stream_handle(self.dev)))
        self._refresh_flips()

    def train_step(self, img, target, weight, lr, scale=None):
        """forward, loss, backward, Adam.  target / weight [B, Ty, Tx, 144].  Returns the
        device loss (no host synchronisation)."""
        out = self.forward(img)
        A = self.activations(*out.shape[:3])
        scale = (1.0 / out.numel()) if scale is None else scale
        self.loss_backward(A, target.contiguous().float(), weight.contiguous().float(), scale)
        self.backward(A)
        self.adam(lr)
        return self.loss


# ---------------------------------------------------------------------- encode / decode
def activated(params, W, H):
    """Per-Gaussian activated quantities of the seven raw tensors (PARAM_NAMES order)."""
    means, scales_raw, rot, colours, phase, opac_raw, curv_raw = params
    mt = torch.tanh(means)
    mx, my = (mt[:, 0] + 1) * 0.5 * W, (mt[:, 1] + 1) * 0.5 * H
    return dict(mx=mx, my=my, scales_raw=scales_raw, rot=rot, colours=colours, phase=phase,
                opac=torch.sigmoid(opac_raw), curv=torch.tanh(curv_raw),
                sx=torch.exp(scales_raw[:, 0]) + 0.1, sy=torch.exp(scales_raw[:, 1]) + 0.1)


def encode_targets(params, W, H):
    """Slot targets [Ty, Tx, 144] and weights of a fitted scene (raw parameters)."""
    a = activated(params, W, H)
    dev = a["mx"].device
    Ty, Tx = -(-H // 16), -(-W // 16)
    tx = torch.clamp((a["mx"] / 16).floor().long(), 0, Tx - 1)
    ty = torch.clamp((a["my"] / 16).floor().long(), 0, Ty - 1)
    tile = ty * Tx + tx
    energy = a["opac"] * (a["colours"] ** 2).sum(1) * a["sx"] * a["sy"]
    o1 = torch.argsort(energy, descending=True, stable=True)
    o2 = torch.argsort(tile[o1], stable=True)
    order = o1[o2]
    ts = tile[order]
    first = torch.searchsorted(ts, torch.arange(Ty * Tx, device=dev))
    rank = torch.arange(order.numel(), device=dev) - first[ts]
    keep = rank < SLOTS
    g, tk, rk = order[keep], ts[keep], rank[keep]
    C = a["colours"].shape[1]
    enc = torch.zeros(g.numel(), SLOT_DIM, device=dev)
    enc[:, 0] = 2 * (a["mx"][g] - 16 * (tk % Tx).float()) / 16 - 1
    enc[:, 1] = 2 * (a["my"][g] - 16 * (tk // Tx).float()) / 16 - 1
    enc[:, 2:4] = a["scales_raw"][g]
    enc[:, 4] = torch.cos(2 * a["rot"][g])
    enc[:, 5] = torch.sin(2 * a["rot"][g])
    for c in range(min(C, 3)):
        amp, ph = a["colours"][g, c], a["phase"][g, c]
        ph = torch.where(amp < 0, ph + math.pi, ph)
        enc[:, 6 + c] = amp.abs()
        enc[:, 9 + 2 * c] = torch.cos(ph)
        enc[:, 10 + 2 * c] = torch.sin(ph)
    enc[:, 15] = a["curv"][g]
    enc[:, 16] = a["opac"][g]
    enc[:, 17] = 1.0
    target = torch.zeros(Ty * Tx, SLOTS, SLOT_DIM, device=dev)
    target[tk, rk] = enc
    weight = torch.zeros_like(target)
    weight[:, :, [6, 7, 8, 16, 17]] = 1.0
    weight[tk, rk] = 1.0
    return target.view(Ty, Tx, OUT_DIM), weight.view(Ty, Tx, OUT_DIM)


class _StraightClamp(torch.autograd.Function):
    """Clamp in the forward, identity in the backward."""

    @staticmethod
    def forward(ctx, x, lo, hi):
        return x.clamp(lo, hi)

    @staticmethod
    def backward(ctx, g):
        return g, None, None


def _clamp(x, lo=None, hi=None):
    """Clamp whose gradient passes through unchanged (straight-through): a slot output
    pushed past a bound still learns, instead of going dead (a colour below 0 renders
    nothing and would never receive a gradient again)."""
    return _StraightClamp.apply(x, lo, hi)


def decode(out, W, H, channels=3, ref=None, min_norm=0.5):
    """Slot encodings [Ty, Tx, 144] -> (seven raw parameter tensors of Ty*Tx*8 Gaussians,
    presence mask [Ty*Tx*8]).  With ``ref`` (a [144] encoding), a slot's phase whose
    predicted (cos, sin) vector is shorter than ``min_norm`` (an uncertain phase) takes
    the reference slot's phase instead.  The clamps pass gradients straight through
    (``_clamp``); the decoded values are the clamped ones."""
    Ty, Tx = out.shape[:2]
    o = out.reshape(Ty * Tx * SLOTS, SLOT_DIM).float()
    t = torch.arange(Ty * Tx, device=out.device).repeat_interleave(SLOTS)
    mx = 16 * (t % Tx).float() + (_clamp(o[:, 0], -1, 1) + 1) * 8
    my = 16 * (t // Tx).float() + (_clamp(o[:, 1], -1, 1) + 1) * 8
    mx = _clamp(mx, 0.01, W - 0.01)
    my = _clamp(my, 0.01, H - 0.01)
    means = torch.stack([torch.atanh(_clamp(2 * mx / W - 1, -0.9999, 0.9999)),
                         torch.atanh(_clamp(2 * my / H - 1, -0.9999, 0.9999))], 1)
    scales = _clamp(o[:, 2:4], -4.0, 3.0)
    rot = 0.5 * torch.atan2(o[:, 5], o[:, 4])
    colours = _clamp(torch.stack([o[:, 6 + c] for c in range(channels)], 1), 0.0)
    phase = torch.stack([torch.atan2(o[:, 10 + 2 * c], o[:, 9 + 2 * c]) for c in range(channels)], 1)
    if ref is not None:
        r = ref.reshape(SLOTS, SLOT_DIM).float().repeat(Ty * Tx, 1)
        for c in range(channels):
            norm = torch.hypot(o[:, 9 + 2 * c], o[:, 10 + 2 * c])
            rp = torch.atan2(r[:, 10 + 2 * c], r[:, 9 + 2 * c])
            phase[:, c] = torch.where(norm < min_norm, rp, phase[:, c])
    phase = torch.remainder(phase, 2 * math.pi)
    curv = torch.atanh(_clamp(o[:, 15], -0.99, 0.99))
    opac = torch.logit(_clamp(o[:, 16], 0.01, 0.99))
    present = o[:, 17] >= 0.5
    return [means.contiguous(), scales.contiguous(), rot.contiguous(), colours.contiguous(),
            phase.contiguous(), opac.contiguous(), curv.contiguous()], present


def network_input(wl):
    """[4, H, W]: the RGB target image and the depth map of a workload."""
    return torch.cat([wl.target_image.float(), wl.target_depth.float()], 0)


def params_from_slots(out, wl, ref=None):
    """Raw parameters of a scene started from slot encodings [Ty, Tx, 144]: the decoded
    Gaussians of the present slots, then the workload's random initial Gaussians up to N."""
    W, H = wl.img_size
    pred, present = decode(out, W, H, wl.gaussians.channels, ref=ref)
    from .model import PARAM_NAMES
    N = wl.gaussians.means_2d.shape[0]
    rest = [getattr(wl.gaussians, n).detach() for n in PARAM_NAMES]
    idx = torch.nonzero(present).flatten()[:N]
    k = idx.numel()
    return [torch.cat([p[idx], r[:N - k]], 0).contiguous() for p, r in zip(pred, rest)]


def all_slot_params(out, wl):
    """Raw parameters of a scene started from slot encodings [Ty, Tx, 144] using every
    slot (the render-tuned network's start): the Ty*Tx*8 decoded Gaussians, then the
    workload's random initial Gaussians up to N."""
    W, H = wl.img_size
    pred, _ = decode(out, W, H, wl.gaussians.channels)
    from .model import PARAM_NAMES
    N = wl.gaussians.means_2d.shape[0]
    rest = [getattr(wl.gaussians, n).detach() for n in PARAM_NAMES]
    k = min(N, pred[0].shape[0])
    return [torch.cat([p[:k], r[:N - k]], 0).contiguous() for p, r in zip(pred, rest)]


def warm_params(net, wl, ref=None):
    """Raw parameters of a warm-started scene: the network's Gaussians (up to 8 per tile,
    the slots it marks present), then the workload's own random initial Gaussians up to
    its count N."""
    return params_from_slots(net.forward(network_input(wl)[None])[0], wl, ref=ref)


def finetune(net, wls, steps, lr=3e-4, log=None, checkpoint=None, every=250):
    """Render-loss fine-tuning: for one scene per step, the network's slots are decoded
    into the scene's first Ty*Tx*8 Gaussians (the rest keep their random start), the
    fused step runs without its optimiser (cvqpg kernels: raster, BLASM, loss, adjoints),
    the raw-parameter gradients flow back through the decode (torch autograd on the
    decode's elementwise maps) into the network's output, and the network takes one Adam
    step on its own kernels.  The objective is the fit's own loss at its first step, so the
    network learns starts that already render close to the targets.  ``checkpoint(k)``,
    when given, is called after every ``every`` steps and after the last one (the caller
    validates and keeps the best network state).  Returns the loss history (every 50
    steps)."""
    from .fused import FusedTrainer
    from .model import PARAM_NAMES
    trs = [FusedTrainer(wl, num_itrs=1000, warmup_iters=0, inject_noise=False) for wl in wls]
    inits = [[getattr(wl.gaussians, n).detach().clone() for n in PARAM_NAMES] for wl in wls]
    imgs = [network_input(wl)[None].contiguous() for wl in wls]
    hist = []
    for k in range(steps):
        i = k % len(trs)
        tr, wl = trs[i], wls[i]
        W, H = wl.img_size
        out = net.forward(imgs[i])
        o = out[0].detach().clone().requires_grad_()
        params, _ = decode(o, W, H, wl.gaussians.channels)
        m = params[0].shape[0]
        with torch.no_grad():
            for p, q, r in zip(tr.params, params, inits[i]):
                p.copy_(torch.cat([q, r[m:]], 0))
        tr.iteration.zero_()
        tr.loss_and_grads()
        gv = tr.grad_views()
        torch.autograd.backward(params, [gv[n][:m] for n in PARAM_NAMES])
        A = net.activations(1, *out.shape[1:3])
        A["grad"]["out"].copy_(o.grad[None])
        net.backward(A)
        # cosine schedule after a 50-step linear warm-up (Adam's first bias-corrected
        # steps move every weight by about lr)
        net.adam(lr * min(1.0, (k + 1) / 50) * 0.5 * (1 + math.cos(math.pi * k / steps)))
        if k % 50 == 0 or k == steps - 1:
            hist.append((k, tr.loss.clone()))
        if checkpoint is not None and ((k + 1) % every == 0 or k == steps - 1):
            checkpoint(k + 1)
    out = [(k, float(v)) for k, v in hist]
    if log:
        log("render fine-tune loss " + " ".join(f"{k}:{v:.4f}" for k, v in out[::max(1, len(out) // 8)]))
    return out


def start_loss(start_fn, wls):
    """Mean loss of the fit's first step (no optimiser step) from ``start_fn``'s starts:
    the validation measure that chooses between candidate starts."""
    from .fused import FusedTrainer
    from .model import PARAM_NAMES
    vals = []
    for wl in wls:
        init = start_fn(wl)
        saved = [getattr(wl.gaussians, n).detach().clone() for n in PARAM_NAMES]
        with torch.no_grad():
            for n, p in zip(PARAM_NAMES, init):
                getattr(wl.gaussians, n).data.copy_(p)
        tr = FusedTrainer(wl, num_itrs=1000, warmup_iters=0, inject_noise=False)
        tr.loss_and_grads()
        vals.append(float(tr.loss))
        with torch.no_grad():
            for n, p in zip(PARAM_NAMES, saved):
                getattr(wl.gaussians, n).data.copy_(p)
        del tr
    return float(np.mean(vals))


# ---------------------------------------------------------------------- experiment
def fit(wl, iters, eval_every, num_itrs=None, init=None, stop_at=None, log=None):
    """Fused fit of one workload from its own parameters (or ``init``), recording the
    per-plane PSNR every ``eval_every`` iterations.  With ``stop_at`` (per-plane PSNR
    targets) it stops at the first evaluation where every plane reaches its target.
    Returns dict(iters, psnr [evals x D], ms (GPU time of the steps), reached)."""
    from .export import plane_metrics
    from .fused import FusedTrainer
    from .model import PARAM_NAMES
    if init is not None:
        with torch.no_grad():
            for n, p in zip(PARAM_NAMES, init):
                getattr(wl.gaussians, n).data.copy_(p)
    tr = FusedTrainer(wl, num_itrs=num_itrs or iters, warmup_iters=min(400, iters // 5))
    W, H = wl.img_size
    tr.step()
    tr.capture()
    hist, it, reached, ms = [], 1, False, 0.0
    target = None if stop_at is None else torch.tensor(stop_at)
    while it < iters:
        n = min(eval_every, iters - it)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(n):
            tr.run()
        b.record()
        it += n
        psnr, _ = plane_metrics(tr.inten, tr.targets, H, W)
        psnr = psnr.cpu()            # one host read per evaluation, in both kinds of fit
        ms += a.elapsed_time(b)      # GPU time of the chunk's steps only
        hist.append(psnr)
        if target is not None and bool((psnr >= target).all()):
            reached = True
            break
        if it % 200 < eval_every and not tr.check_capacity():
            tr.grow()
    return dict(iters=it, psnr=torch.stack(hist), ms=ms, reached=reached,
                ms_per_step=ms / max(1, it - 1), trainer=tr)


def make_workload(spec, seed, device):
    from .train import Workload
    return Workload(f"warm{seed}", device=device, seed=seed, spec=spec)


def scene_spec(W, H, density=150_000 / (1920 * 1080), num_planes=4, d_val=3e-3,
               volume_depth=6e-3):
    """The 1080p recipe (RGB, 8 um, 4 planes over 0-6 mm, 0.0723 Gaussians per pixel) at
    another frame size."""
    return dict(W=W, H=H, N=int(round(density * W * H)), wavelengths=[639e-9, 515e-9, 473e-9],
                pitch=8e-6, num_planes=num_planes, d_val=d_val, volume_depth=volume_depth,
                kappa=0.02)


def crop_specs(size, source, n, seed=0):
    """Specs of n training windows of ``size`` cut at seeded offsets from frames of the
    ``source`` size (4 windows per frame), so that training sees the target and depth
    statistics of the large frame at the cost of fitting the small one.  Returns
    (specs, seeds): window i uses frame seed seeds[i]."""
    W, H = size
    Ws, Hs = source
    rng = np.random.default_rng(seed + 31)
    specs, seeds = [], []
    for i in range(n):
        sp = scene_spec(W, H)
        sp["crop"] = (int(rng.integers(0, Ws - W + 1)), int(rng.integers(0, Hs - H + 1)), Ws, Hs)
        specs.append(sp)
        seeds.append(40_000 + seed * 100_000 + i // 4)
    return specs, seeds


def make_pairs(spec, seeds, iters, device, batch=16, log=None):
    """Training pairs from plain fits: inputs [n, 4, H, W], slot targets and weights.
    ``spec`` is one spec for every seed or a list of specs (one per seed)."""
    from .batch import BatchTrainer
    from .fused import FusedTrainer
    specs = spec if isinstance(spec, (list, tuple)) else [spec] * len(seeds)
    xs, ts, ws = [], [], []
    for i in range(0, len(seeds), batch):
        wls = [make_workload(sp, s, device) for sp, s in zip(specs[i:i + batch], seeds[i:i + batch])]
        trs = [FusedTrainer(wl, num_itrs=iters, warmup_iters=min(400, iters // 5)) for wl in wls]
        bt = BatchTrainer(None, trainers=trs, streams=4)
        bt.capture()
        for it in range(iters):
            bt.run()
            if (it + 1) % 200 == 0 and not bt.check_capacity():
                bt.grow()
                bt.capture()
        W, H = specs[i]["W"], specs[i]["H"]
        for wl, tr in zip(wls, trs):
            t, w = encode_targets(tr.params, W, H)
            xs.append(network_input(wl))
            ts.append(t)
            ws.append(w)
        if log:
            log(f"pairs {i + len(wls)}/{len(seeds)}: losses "
                + " ".join(f"{float(t.loss):.4f}" for t in trs[:4]))
        del bt, trs, wls
    return torch.stack(xs), torch.stack(ts), torch.stack(ws)


def train(net, xs, ts, ws, steps, batch=16, lr=1e-3, seed=0, log=None):
    """Adam on random minibatches with a cosine learning rate; returns the loss history."""
    g = torch.Generator().manual_seed(seed)
    n = xs.shape[0]
    hist = []
    for k in range(steps):
        idx = torch.randint(0, n, (min(batch, n),), generator=g).to(xs.device)
        lr_k = lr * 0.5 * (1 + math.cos(math.pi * k / steps))
        loss = net.train_step(xs[idx], ts[idx], ws[idx], lr_k)
        if k % 50 == 0 or k == steps - 1:
            hist.append((k, loss.clone()))
        if log and k % 250 == 249:
            log(f"train step {k + 1}/{steps}: loss {float(loss):.4f}")
    out = [(k, float(v)) for k, v in hist]
    if log:
        log("train loss " + " ".join(f"{k}:{v:.4f}" for k, v in out[::max(1, len(out) // 8)]))
    return out


def time_to_psnr(spec, seeds, iters, eval_every, device, starts, tol_db=0.1, log=None):
    """Per held-out scene: plain fit for ``iters`` iterations -> final per-plane PSNR P;
    then, for each named start in ``starts`` (name -> function(workload) -> raw
    parameters), a fit from that start until every plane reaches P - tol_db (at most
    ``iters`` iterations).  Times are GPU milliseconds of the fit steps (same step kernel
    and N for every start); a start's own cost (network forward, decode) is added to it.
    Returns {name: [row per scene]}."""
    out = {name: [] for name in starts}
    for s in seeds:
        wl = make_workload(spec, s, device)
        plain = fit(wl, iters, eval_every)
        final = plain["psnr"][-1]
        need = (final - tol_db).tolist()
        first_hit = next((i for i, p in enumerate(plain["psnr"]) if bool((p >= final - tol_db).all())),
                         len(plain["psnr"]) - 1)
        del wl
        for name, init_fn in starts.items():
            wl2 = make_workload(spec, s, device)
            torch.cuda.synchronize()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            init = init_fn(wl2)
            b.record()
            warm = fit(wl2, iters, eval_every, num_itrs=iters, init=init, stop_at=need)
            t_net = a.elapsed_time(b)
            row = dict(seed=s, plain_final_psnr=[round(float(v), 3) for v in final],
                       plain_iters=plain["iters"], plain_ms=plain["ms"],
                       plain_iters_to_target=1 + (first_hit + 1) * eval_every,
                       warm_iters=warm["iters"], warm_ms=warm["ms"] + t_net, warm_net_ms=t_net,
                       warm_reached=warm["reached"],
                       psnr_first_eval_plain=[round(float(v), 3) for v in plain["psnr"][0]],
                       psnr_first_eval_warm=[round(float(v), 3) for v in warm["psnr"][0]])
            out[name].append(row)
            if log:
                log(f"seed {s} [{name}]: plain {plain['iters']} it {plain['ms']:.0f} ms -> "
                    f"{row['plain_final_psnr']}; start {warm['iters']} it {row['warm_ms']:.0f} ms "
                    f"reached={warm['reached']}; first-eval PSNR plain "
                    f"{row['psnr_first_eval_plain']} start {row['psnr_first_eval_warm']}")
            del warm, wl2
        del plain
    return out


def summarize(rows):
    reached = [r for r in rows if r["warm_reached"]]
    ratio = [r["warm_ms"] / r["plain_ms"] for r in rows]
    return dict(scenes=len(rows), warm_reached=len(reached),
                median_time_ratio=float(np.median(ratio)) if ratio else None,
                mean_time_ratio=float(np.mean(ratio)) if ratio else None,
                median_warm_iters=float(np.median([r["warm_iters"] for r in rows])),
                plain_iters=rows[0]["plain_iters"] if rows else None)


def mean_slots(ts, ws):
    """Image-independent slot encoding [144]: the weighted mean of the training targets
    over pairs and tiles (broadcast to any tile grid)."""
    t = ts.reshape(-1, OUT_DIM)
    w = ws.reshape(-1, OUT_DIM)
    return (t * w).sum(0) / w.sum(0).clamp(min=1)


def experiment(device="cuda", train_size=(1920, 1080), n_train=16, fit_iters=600,
               train_steps=600, ft_size=(512, 512), ft_source=(1920, 1080), ft_scenes=16,
               ft_steps=1500, ft_lr=3e-4, val_scenes=4,
               eval_sets=(((1920, 1080), 4, 2000), ((512, 512), 4, 1000)),
               eval_every=25, seed=0, log=None, train_source=None, with_mean=True):
    """The whole warm-start study on one GPU:

    1. training pairs from plain fits (whole ``train_size`` frames, or windows of it cut
       from ``train_source`` frames);
    2. supervised training on the pairs' slot encodings, from a head biased to their mean;
    3. render-loss fine-tuning (``finetune``) on ``ft_scenes`` scenes of ``ft_size``
       (windows cut from ``ft_source`` frames, or whole frames when it is None);
    4. validation: the first-step loss on ``val_scenes`` further scenes picks the best
       fine-tuning checkpoint (every 250 steps) and then chooses between it (every slot)
       and the supervised network (present slots, gated phases); the plain start stays
       the fallback and the yardstick;
    5. time-to-PSNR on held-out scenes of the chosen warm start against the plain start,
       and of the image-independent start from the mean slot encoding, which isolates
       what the network's conditioning on the image adds.
    Returns a dict."""
    import time as _time
    dev = torch.device(device)
    t0 = _time.perf_counter()
    if train_source is None:
        spec = scene_spec(*train_size)
        seeds = [10_000 + seed * 100_000 + i for i in range(n_train)]
    else:
        spec, seeds = crop_specs(train_size, train_source, n_train, seed)
    xs, ts, ws = make_pairs(spec, seeds, fit_iters, dev, log=log)
    torch.cuda.synchronize(dev)
    t_pairs = _time.perf_counter() - t0
    tm = mean_slots(ts, ws)
    net = WarmStartNet(dev, seed=seed)
    net.set_output_bias(tm)   # the network starts as the image-blind mean predictor
    t1 = _time.perf_counter()
    hist = train(net, xs, ts, ws, train_steps, lr=1e-3, seed=seed, log=log)
    torch.cuda.synchronize(dev)
    t_train = _time.perf_counter() - t1
    del xs, ts, ws
    sup = WarmStartNet(dev, seed=seed)
    sup.load_state_dict(net.state_dict())

    def scenes(n, salt):
        if ft_source is None:
            sp = scene_spec(*ft_size)
            return [make_workload(sp, 30_000 + seed * 100_000 + salt * 1000 + i, dev)
                    for i in range(n)]
        sps, sds = crop_specs(ft_size, ft_source, n, seed + salt)
        return [make_workload(sp, s, dev) for sp, s in zip(sps, sds)]

    val = scenes(val_scenes, 11)
    tuned = lambda wl: all_slot_params(net.forward(network_input(wl)[None])[0], wl)
    gated = lambda wl: warm_params(sup, wl, ref=tm)
    # the render loss is not monotone over fine-tuning (a step on one scene can undo
    # others): the network is validated every 250 steps and the best state is kept
    best = dict(loss=float("inf"), state=None, step=None)

    def checkpoint(k):
        lv = start_loss(tuned, val)
        if lv < best["loss"]:
            best.update(loss=lv, state=net.state_dict(), step=k)
        if log:
            log(f"fine-tune step {k}: validation first-step loss {lv:.4f}")

    t2 = _time.perf_counter()
    ft_hist = (finetune(net, scenes(ft_scenes, 7), ft_steps, ft_lr, log=log,
                        checkpoint=checkpoint) if ft_steps else [])
    if best["state"] is not None:
        net.load_state_dict(best["state"])
    torch.cuda.synchronize(dev)
    t_ft = _time.perf_counter() - t2
    loss_tuned = best["loss"]
    loss_gated = start_loss(gated, val)
    choice = "fine-tuned" if loss_tuned <= loss_gated else "supervised"
    if log:
        log(f"validation first-step loss: fine-tuned {loss_tuned:.4f}, supervised {loss_gated:.4f}"
            f" -> {choice}")
    del val
    result = dict(network=dict(params=net.n_params, train_pairs=n_train,
                               train_frame=list(train_size),
                               train_source=list(train_source) if train_source else None,
                               fit_iters=fit_iters, train_steps=train_steps,
                               final_train_loss=hist[-1][1], ft_frame=list(ft_size),
                               ft_scenes=ft_scenes, ft_steps=ft_steps,
                               ft_loss_first=ft_hist[0][1] if ft_hist else None,
                               ft_loss_last=ft_hist[-1][1] if ft_hist else None,
                               val_loss_finetuned=loss_tuned, ft_best_step=best["step"],
                               val_loss_supervised=loss_gated,
                               chosen=choice, pairs_s=t_pairs, train_s=t_train,
                               finetune_s=t_ft),
                  evals=[])
    grid = lambda wl: tm.expand(-(-wl.img_size[1] // 16), -(-wl.img_size[0] // 16), OUT_DIM)
    starts = {"net": tuned if choice == "fine-tuned" else gated}
    if with_mean:
        starts["mean"] = lambda wl: params_from_slots(grid(wl), wl)
    for (W, H), n, iters in eval_sets:
        espec = scene_spec(W, H)
        seeds = [20_000 + seed * 100_000 + i for i in range(n)]
        rows = time_to_psnr(espec, seeds, iters, eval_every, dev, starts, log=log)
        result["evals"].append(dict(
            frame=[W, H], gaussians=espec["N"], plain_iters=iters, net=summarize(rows["net"]),
            mean_start=summarize(rows["mean"]) if with_mean else None, rows=rows["net"]))
    return result
