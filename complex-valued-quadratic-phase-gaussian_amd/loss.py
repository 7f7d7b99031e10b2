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

These are PyTorch implementations (setup work and the PyTorch baseline step); the
fused training step evaluates the same loss in the gfx950 loss kernels.
"""
import math

import torch
import torch.nn.functional as F

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


def GaussianLoss(pred, target, lambda_ssim=0.025):  # noqa: N802 (reference name)
    return lambda_ssim * (1 - ssim(pred, target)) * 0.2


GaussianLossGrayscale = GaussianLoss


class multiplane_loss_odak:  # noqa: N801 (reference class name)
    def __init__(self, target_image, target_depth, blur_ratio=0.25, target_blur_size=10,
                 number_of_planes=4, weights=(1.0, 2.1, 0.6), multiplier=1.0, scheme="defocus",
                 reduction="mean", split_ratio=1.0, device=torch.device("cpu")):
        self.device = torch.device(device)
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
        """utils/propagator.py:392-415."""
        L = self.target_blur_size
        P = self.number_of_planes
        for ch in range(self.target_image.shape[0]):
            cache = self.targets[:, ch].clone()
            total = cache.sum(0)
            for i in range(P):
                acc = torch.zeros_like(cache[i])
                for j in range(P):
                    if torch.sum(cache[j]) > 0:
                        s = 0.0 if i == j else float(int(abs(i - j) * self.blur_ratio))
                        ker = odak_gaussian_2d(L, s, self.device)
                        ker = (ker / ker.sum())[None, None]
                        bl = F.conv2d(total[None, None], ker, padding="same")[0, 0]
                        acc = acc + bl * self.masks[j, ch].abs()
                self.targets[i, ch] = acc
        self.targets = self.targets.detach().clone() * self.multiplier

    def __call__(self, image, target, plane_id=None, inject_noise=False, noise_ratio=1e-3, noise=None):
        """utils/propagator.py:417-449; ``noise`` (optional) replaces the internal
        randn draw so a caller can pin the perturbation."""
        l2 = self.weights[0] * F.mse_loss(image, target)
        mask = self.masks if plane_id is None else self.masks[plane_id]
        if inject_noise:
            if noise is None:
                noise = torch.randn_like(target)
            target = target + noise * noise_ratio * (target.max() - target.min())
        l2_mask = self.weights[1] * F.mse_loss(image * mask, target * mask)
        l2_cor = self.weights[2] * F.mse_loss(image * target, target * target)
        return l2 + l2_mask + l2_cor


def multiplane_loss(target_image, target_depth, args_prop):
    """utils/data_utils.py:66-82."""
    fn = multiplane_loss_odak(target_image=target_image, target_depth=target_depth,
                              target_blur_size=20, number_of_planes=args_prop.num_planes,
                              blur_ratio=8, weights=[1.0, 1.0, 1.0, 0.0], scheme="defocus",
                              reduction="mean", split_ratio=getattr(args_prop, "split_ratio", 1.0),
                              device=target_image.device)
    targets, mask, _ = fn.get_targets()
    return targets, fn, mask
