"""Multi-plane defocus loss with the reference's interface.

* ``multiplane_loss_odak`` -- utils/propagator.py:293-449: depth slicing into plane
  masks (set_targets), defocus-blurred per-plane targets (add_defocus_blur, odak
  Gaussian kernel on a linspace(-L/2, L/2, L) grid, sigma 0 -> 1e-5) and the three-term
  masked L2 ``__call__(image, target, plane_id, inject_noise, noise_ratio)``.
* ``GaussianLoss`` / ``GaussianLossGrayscale`` -- utils/data_utils.py:22-63:
  0.025 * 0.2 * (1 - SSIM), SSIM with an 11-tap sigma 1.5 Gaussian window, K = (0.01,
  0.03), data range 1, valid convolution, mean over map and channels (the published
  algorithm of pytorch-msssim 1.0.0, which is not installed here).
* ``multiplane_loss`` -- factory of utils/data_utils.py:66-82.

``multiplane_loss_odak.__call__`` and ``GaussianLoss`` run on the gfx950 kernels
(``cvqpg_plane_l2``, ``cvqpg_ssim_term``, differentiable through torch.autograd.Function)
when ``backend='hip'`` (the default) and in PyTorch when ``backend='torch'`` (the
reference-structured baseline step).  There is no fallback between them: the HIP backend
raises if the library is missing or the tensors are not on the GPU.  Target synthesis
(``set_targets`` / ``add_defocus_blur``, once per fit) stays in PyTorch.  The fused
training step evaluates the whole multi-plane loss in one pass of the loss kernel.
"""
import math

import torch
import torch.nn.functional as F

from . import _lib

BACKEND = "hip"


def _need_gpu(*ts):
    for t in ts:
        if not t.is_cuda:
            raise RuntimeError("the HIP loss kernels need GPU tensors (use backend='torch' on "
                               "the CPU)")


class _PlaneL2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, target, mask, noise, noise_ratio, weights):
        _need_gpu(image, target, mask)
        im = image.detach().contiguous().float()
        tg = target.detach().contiguous().float()
        mk = mask.detach().contiguous().float()
        if not (im.numel() == tg.numel() == mk.numel()):
            raise ValueError(f"image, target and mask must have the same number of elements "
                             f"({im.numel()}, {tg.numel()}, {mk.numel()})")
        nz = None if noise is None else noise.detach().contiguous().float()
        L = _lib.lib()
        n = im.numel()
        ws = torch.empty(int(L.cvqpg_plane_l2_workspace_size(n)), dtype=torch.uint8,
                         device=im.device)
        grad = torch.empty_like(im) if image.requires_grad else None
        out = torch.empty(1, device=im.device)
        _lib.check(L.cvqpg_plane_l2(_lib.ptr(im), _lib.ptr(tg), _lib.ptr(mk), _lib.ptr(nz),
                                    float(noise_ratio), float(weights[0]), float(weights[1]),
                                    float(weights[2]), n, _lib.ptr(ws), ws.numel(),
                                    _lib.ptr(grad), _lib.ptr(out),
                                    _lib.stream_handle(im.device)))
        ctx.save_for_backward(grad if grad is not None else out)
        ctx.has_grad = grad is not None
        ctx.shape = image.shape
        return out[0]

    @staticmethod
    def backward(ctx, g):
        if not ctx.has_grad:
            return None, None, None, None, None, None
        (grad,) = ctx.saved_tensors
        return (grad * g).view(ctx.shape), None, None, None, None, None


class _SSIMTerm(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pred, target):
        _need_gpu(pred, target)
        shape = pred.shape
        if pred.dim() == 4:
            if pred.shape[0] != 1:
                raise ValueError("the HIP SSIM term takes one image [C, H, W] or [1, C, H, W]")
            pred, target = pred[0], target[0]
        im = pred.detach().contiguous().float()
        tg = target.detach().contiguous().float()
        C, H, W = im.shape
        L = _lib.lib()
        nb = int(L.cvqpg_loss_workspace_size(C, 1, H, W))
        if nb == 0:
            raise ValueError(f"image {H}x{W} is smaller than the 11x11 SSIM window")
        ws = torch.empty(nb, dtype=torch.uint8, device=im.device)
        grad = torch.empty_like(im)
        out = torch.empty(1, device=im.device)
        _lib.check(L.cvqpg_ssim_term(_lib.ptr(im), _lib.ptr(tg), C, H, W, _lib.ptr(ws),
                                     ws.numel(), _lib.ptr(grad), _lib.ptr(out),
                                     _lib.stream_handle(im.device)))
        ctx.save_for_backward(grad)
        ctx.shape = shape
        return out[0]

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return (grad * g).view(ctx.shape), None

SSIM_WIN, SSIM_SIGMA, SSIM_K = 11, 1.5, (0.01, 0.03)


def odak_gaussian_2d(length, sigma, device="cpu"):
    x = torch.linspace(-length / 2.0, length / 2.0, length, device=device)
    X, Y = torch.meshgrid(x, x, indexing="ij")
    s = sigma if sigma != 0 else 1e-5
    return 1.0 / (2.0 * math.pi * s * s) * torch.exp(-(X ** 2 / (2.0 * s ** 2) + Y ** 2 / (2.0 * s ** 2)))


def gauss_window(device, dtype=torch.float32):
    c = torch.arange(SSIM_WIN, dtype=dtype, device=device) - SSIM_WIN // 2
    g = torch.exp(-(c ** 2) / (2 * SSIM_SIGMA ** 2))
    return g / g.sum()


def _conv_same(img, ker):
    """2-D correlation of img [H, W] with the odd L x L kernel ker under zero padding (what
    F.conv2d(img[None, None], ker[None, None], padding="same") computes), as a sum of
    L^2 shifted copies: the library convolution picks a naive direct kernel for this
    shape on some devices (hundreds of ms per 1080p plane)."""
    L = ker.shape[0]
    r = L // 2
    H, W = img.shape
    pad = F.pad(img[None, None], (r, r, r, r))[0, 0]
    kv = ker.detach().cpu().tolist()
    out = torch.zeros_like(img)
    for a in range(L):
        for b in range(L):
            out.add_(pad[a:a + H, b:b + W], alpha=kv[a][b])
    return out


def _gfilter(x, w):
    C, n = x.shape[1], w.numel()
    x = F.conv2d(x, w.view(1, 1, 1, n).expand(C, 1, 1, n), groups=C)
    return F.conv2d(x, w.view(1, 1, n, 1).expand(C, 1, n, 1), groups=C)


def ssim(X, Y, data_range=1.0):
    if X.dim() == 3:
        X, Y = X.unsqueeze(0), Y.unsqueeze(0)
    w = gauss_window(X.device, X.dtype)
    C1, C2 = (SSIM_K[0] * data_range) ** 2, (SSIM_K[1] * data_range) ** 2
    mu1, mu2 = _gfilter(X, w), _gfilter(Y, w)
    s11 = _gfilter(X * X, w) - mu1 ** 2
    s22 = _gfilter(Y * Y, w) - mu2 ** 2
    s12 = _gfilter(X * Y, w) - mu1 * mu2
    cs = (2 * s12 + C2) / (s11 + s22 + C2)
    smap = ((2 * mu1 * mu2 + C1) / (mu1 ** 2 + mu2 ** 2 + C1)) * cs
    return smap.flatten(2).mean(-1).mean()


def GaussianLoss(pred, target, lambda_ssim=0.025, backend=None):  # noqa: N802 (reference name)
    """utils/data_utils.py:22-41: lambda_ssim * (1 - SSIM(pred, target)) * 0.2."""
    if (backend or BACKEND) == "hip":
        if lambda_ssim != 0.025:
            raise ValueError("the HIP SSIM term uses the reference's lambda_ssim = 0.025")
        return _SSIMTerm.apply(pred, target)
    return lambda_ssim * (1 - ssim(pred, target)) * 0.2


GaussianLossGrayscale = GaussianLoss


class multiplane_loss_odak:  # noqa: N801 (reference class name)
    def __init__(self, target_image, target_depth, blur_ratio=0.25, target_blur_size=10,
                 number_of_planes=4, weights=(1.0, 2.1, 0.6), multiplier=1.0, scheme="defocus",
                 reduction="mean", split_ratio=1.0, device=torch.device("cpu"), backend=None):
        self.device = torch.device(device)
        self.backend = backend or BACKEND
        self._draws = 0
        self.target_image = target_image.float().to(self.device)
        self.target_depth = target_depth.float().to(self.device)
        self.target_blur_size = target_blur_size + (1 if target_blur_size % 2 == 0 else 0)
        self.number_of_planes = number_of_planes
        self.multiplier = multiplier
        self.weights = list(weights)
        self.reduction = reduction
        self.blur_ratio = blur_ratio
        self.split_ratio = split_ratio
        self.set_targets()
        if scheme == "defocus":
            self.add_defocus_blur()

    def get_targets(self):
        div = max(self.number_of_planes - 1, 1)
        return self.targets.clone(), self.masks.clone(), self.target_depth.clone() / div

    def set_targets(self):
        """utils/propagator.py:361-390."""
        ratio = self.split_ratio if self.number_of_planes == 2 else 1.0
        self.target_depth = torch.round(torch.pow(self.target_depth, ratio) * (self.number_of_planes - 1))
        P = self.number_of_planes
        C, H, W = self.target_image.shape
        self.targets = torch.zeros(P, C, H, W, device=self.device)
        self.masks = torch.zeros_like(self.targets)
        for i in range(P):
            m = (self.target_depth == i).float().reshape(-1, H, W)[0]
            for ch in range(C):
                self.targets[i, ch] = self.target_image[ch] * m
                self.masks[i, ch] = m

    def add_defocus_blur(self):
        """utils/propagator.py:392-415.  The blurred image depends only on the channel and
        the blur level s = int(|i - j| blur_ratio), so each (channel, s) is filtered once
        and reused for every plane pair with that level."""
        L = self.target_blur_size
        P = self.number_of_planes
        for ch in range(self.target_image.shape[0]):
            cache = self.targets[:, ch].clone()
            total = cache.sum(0)
            blurred = {}
            for i in range(P):
                acc = torch.zeros_like(cache[i])
                for j in range(P):
                    if torch.sum(cache[j]) > 0:
                        s = 0.0 if i == j else float(int(abs(i - j) * self.blur_ratio))
                        if s not in blurred:
                            ker = odak_gaussian_2d(L, s, self.device)
                            blurred[s] = _conv_same(total, ker / ker.sum())
                        acc = acc + blurred[s] * self.masks[j, ch].abs()
                self.targets[i, ch] = acc
        self.targets = self.targets.detach().clone() * self.multiplier

    def __call__(self, image, target, plane_id=None, inject_noise=False, noise_ratio=1e-3, noise=None):
        """utils/propagator.py:417-449; ``noise`` (optional) replaces the internal
        randn draw so a caller can pin the perturbation.  With the HIP backend the draw
        is the Philox4x32-10 normal of the loss kernels (cvqpg_loss_noise), keyed by a
        per-object call counter."""
        mask = self.masks if plane_id is None else self.masks[plane_id]
        if self.backend == "hip":
            if inject_noise and noise is None:
                noise = self._draw(target)
            return _PlaneL2.apply(image, target, mask, noise if inject_noise else None,
                                  noise_ratio, self.weights[:3])
        l2 = self.weights[0] * F.mse_loss(image, target)
        if inject_noise:
            if noise is None:
                noise = torch.randn_like(target)
            target = target + noise * noise_ratio * (target.max() - target.min())
        l2_mask = self.weights[1] * F.mse_loss(image * mask, target * mask)
        l2_cor = self.weights[2] * F.mse_loss(image * target, target * target)
        return l2 + l2_mask + l2_cor

    def _draw(self, target):
        """Standard normals shaped like ``target`` from the loss kernels' generator."""
        _need_gpu(target)
        t = target.reshape(-1, target.shape[-2], target.shape[-1])
        C = 3 if t.shape[0] % 3 == 0 else 1
        planes = t.shape[0] // C
        if planes > 16:
            raise ValueError(f"cannot draw noise shaped {tuple(target.shape)}")
        out = torch.empty(t.shape, device=target.device)
        self._draws += 1
        _lib.check(_lib.lib().cvqpg_loss_noise(0x5EED + self._draws, None, C, planes,
                                               t.shape[1], t.shape[2], _lib.ptr(out),
                                               _lib.stream_handle(target.device)))
        return out.view_as(target)


def multiplane_loss(target_image, target_depth, args_prop):
    """utils/data_utils.py:66-82."""
    fn = multiplane_loss_odak(target_image=target_image, target_depth=target_depth,
                              target_blur_size=20, number_of_planes=args_prop.num_planes,
                              blur_ratio=8, weights=[1.0, 1.0, 1.0, 0.0], scheme="defocus",
                              reduction="mean", split_ratio=getattr(args_prop, "split_ratio", 1.0),
                              device=target_image.device)
    targets, mask, _ = fn.get_targets()
    return targets, fn, mask
