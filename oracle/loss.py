"""Multi-plane defocus loss and target synthesis, restated in torch (dtype-generic).

Restates
  * ``multiplane_loss_odak.__call__``       utils/propagator.py:417-449
      l2      = MSE(I, T)
      l2_mask = MSE(I * M_d, T~ * M_d)        T~ = T + noise (inject_noise=True)
      l2_cor  = MSE(I * T~, T~ * T~)
    with weights [1, 1, 1] (utils/data_utils.py:75).
  * ``GaussianLoss`` / ``GaussianLossGrayscale`` utils/data_utils.py:22-63:
      0.025 * 0.2 * (1 - SSIM(I, T)), SSIM from pytorch-msssim 1.0.0 (not
      vendored, not installable here; restated from its published algorithm:
      Gaussian window 11 taps, sigma 1.5, K = (0.01, 0.03), data_range 1,
      separable 'valid' convolution, mean over the map then over channels).
      Parity against pytorch-msssim itself is unpinned; the textbook SSIM of
      Wang et al. 2004 pins it.
  * the per-step loss assembly  train_2d_gaussian_rgb.py:191-197
  * ``set_targets`` / ``add_defocus_blur``  utils/propagator.py:361-415, with the
    factory values of utils/data_utils.py:66-82 (blur size 20 -> 21, blur ratio 8)
    and odak's ``generate_2d_gaussian`` (kernel grid linspace(-L/2, L/2, L),
    sigma 0 -> 1e-5; recalled, parity unpinned).

Pinned by (tests/test_oracle_loss.py): MSE closed forms, SSIM(x, x) = 1, the
separable filter against a brute-force 11x11 window sum in pure Python on a
tiny image, the SSIM of a constant pair against its closed form, and the
gradient against FP64 central finite differences.
"""
import math

import torch
import torch.nn.functional as F

SSIM_WIN = 11
SSIM_SIGMA = 1.5
SSIM_K = (0.01, 0.03)
LAMBDA_SSIM = 0.025


def gauss_1d(size=SSIM_WIN, sigma=SSIM_SIGMA, dtype=torch.float64):
    coords = torch.arange(size, dtype=dtype) - size // 2
    g = torch.exp(-(coords ** 2) / (2 * sigma ** 2))
    return g / g.sum()


def gfilter(x, win):
    """Separable 'valid' Gaussian filter over the last two dims of [B, C, H, W]."""
    C = x.shape[1]
    n = win.numel()
    w = win.to(x.dtype)
    out = F.conv2d(x, w.view(1, 1, 1, n).expand(C, 1, 1, n), groups=C)
    out = F.conv2d(out, w.view(1, 1, n, 1).expand(C, 1, n, 1), groups=C)
    return out


def ssim(X, Y, data_range=1.0):
    """Mean SSIM over the valid map and channels.  X, Y: [C, H, W] or [B, C, H, W]."""
    if X.dim() == 3:
        X, Y = X.unsqueeze(0), Y.unsqueeze(0)
    win = gauss_1d(dtype=X.dtype)
    C1 = (SSIM_K[0] * data_range) ** 2
    C2 = (SSIM_K[1] * data_range) ** 2
    mu1, mu2 = gfilter(X, win), gfilter(Y, win)
    s11 = gfilter(X * X, win) - mu1 ** 2
    s22 = gfilter(Y * Y, win) - mu2 ** 2
    s12 = gfilter(X * Y, win) - mu1 * mu2
    cs = (2 * s12 + C2) / (s11 + s22 + C2)
    smap = ((2 * mu1 * mu2 + C1) / (mu1 ** 2 + mu2 ** 2 + C1)) * cs
    return smap.flatten(2).mean(-1).mean()


def ssim_term(pred, target):
    return LAMBDA_SSIM * (1 - ssim(pred, target)) * 0.2


def plane_loss(pred, target, mask, noise=None):
    """multiplane_loss_odak.__call__ with weights [1,1,1]; ``noise`` is the already
    scaled perturbation (randn * 1e-3 * (max T - min T)) or None."""
    l2 = F.mse_loss(pred, target)
    tn = target if noise is None else target + noise
    l2_mask = F.mse_loss(pred * mask, tn * mask)
    l2_cor = F.mse_loss(pred * tn, tn * tn)
    return l2 + l2_mask + l2_cor


def step_loss(intens, targets, masks, H, W, noises=None):
    """train:191-197: intens [D, C, pH, pW] -> sum over planes of plane_loss + SSIM term."""
    from .optics import crop_center
    loss = 0.0
    for d in range(intens.shape[0]):
        pred = torch.clamp(crop_center(intens[d], (H, W)), 0.0, 1.0)
        loss = loss + plane_loss(pred, targets[d], masks[d], None if noises is None else noises[d])
        loss = loss + ssim_term(pred, targets[d])
    return loss


def odak_gaussian_2d(length, sigma, dtype=torch.float64):
    x = torch.linspace(-length / 2.0, length / 2.0, length, dtype=dtype)
    X, Y = torch.meshgrid(x, x, indexing="ij")
    s = sigma if sigma != 0 else 1e-5
    return 1.0 / (2.0 * math.pi * s * s) * torch.exp(-(X ** 2 / (2.0 * s ** 2) + Y ** 2 / (2.0 * s ** 2)))


def make_targets(image, depth, num_planes, blur_size=20, blur_ratio=8, split_ratio=1.0):
    """set_targets + add_defocus_blur (propagator.py:361-415).  image [C,H,W], depth [1,H,W].
    Returns (targets [D,C,H,W], masks [D,C,H,W], quantised depth)."""
    if blur_size % 2 == 0:
        blur_size += 1
    ratio = split_ratio if num_planes == 2 else 1.0
    q = torch.round(torch.pow(depth, ratio) * (num_planes - 1))
    C, H, W = image.shape
    targets = torch.zeros(num_planes, C, H, W, dtype=image.dtype)
    masks = torch.zeros_like(targets)
    for i in range(num_planes):
        m = (q == i).to(image.dtype)
        for ch in range(C):
            targets[i, ch] = image[ch] * m[0]
            masks[i, ch] = m[0]
    out = torch.zeros_like(targets)
    for ch in range(C):
        cache = targets[:, ch].clone()
        total = cache.sum(0)
        for i in range(num_planes):
            acc = torch.zeros(H, W, dtype=image.dtype)
            for j in range(num_planes):
                if torch.sum(cache[j]) > 0:
                    s = 0.0 if i == j else float(int(abs(i - j) * blur_ratio))
                    ker = odak_gaussian_2d(blur_size, s, image.dtype)
                    ker = ker / ker.sum()
                    bl = F.conv2d(total[None, None], ker[None, None], padding="same")[0, 0]
                    acc = acc + bl * masks[j, ch].abs()
            out[i, ch] = acc
    return out, masks, q
