"""Phase-only hologram export and evaluation metrics, restated in numpy (FP64).

Test infrastructure only (see the package docstring).

Restates
  * ``dpac`` and the 8-bit quantisation of ``export_poh.py:110-119,173-178``:
      phase = angle(h) mod 2 pi, A = clamp(|h|, 0, 1), offset = arccos(A),
      (even, even) and (odd, odd) pixels take phase - offset, the other two phase +
      offset, the result mod 2 pi; level = round(. / 2 pi * 256) mod 256 with
      torch.round's half-to-even rule.
  * ``evaluate`` PSNR and SSIM of ``train_2d_gaussian_rgb.py:98-112``, which call
    scikit-image 0.2x ``peak_signal_noise_ratio`` and ``structural_similarity``
    (data_range 1, channel_axis 2).  scikit-image is not installed here; its published
    algorithm is restated: PSNR = 10 log10(1 / MSE) over all pixels and channels;
    SSIM per channel with a 7 x 7 uniform window (scipy.ndimage.uniform_filter, as
    scikit-image uses), sample covariance (factor 49 / 48), C1 = 0.01^2, C2 = 0.03^2,
    the mean of the SSIM map over the interior 3 pixels away from each border, then
    the mean over the channels.  Pinned by brute-force windows on small images,
    SSIM(x, x) = 1 and the PSNR closed form; parity with scikit-image itself is
    unpinned.
"""
import numpy as np
from scipy import ndimage

TWO_PI = 2.0 * np.pi


def dpac(phase, amplitude):
    """export_poh.py:110-119 (arrays [..., H, W])."""
    off = np.arccos(np.clip(amplitude, 0.0, 1.0))
    lo, hi = phase - off, phase + off
    yy, xx = np.indices(phase.shape[-2:])
    odd = ((yy + xx) & 1).astype(bool)
    return np.mod(np.where(odd, hi, lo), TWO_PI)


def quantize(v):
    """torch.round (half to even) then mod 256 (export_poh.py:178)."""
    return (np.rint(v).astype(np.int64) % 256).astype(np.uint8)


def dpac_levels(field):
    """8-bit DPAC levels of a complex field [..., H, W] (export_poh.py:173-178) and the
    unrounded level value (for tie-aware comparisons)."""
    phase = np.mod(np.angle(field), TWO_PI)
    amp = np.clip(np.abs(field), 0.0, 1.0)
    v = dpac(phase, amp) / TWO_PI * 256.0
    return quantize(v), v


def crop_center(x, H, W):
    """odak crop_center on the last two axes (even sizes)."""
    r0 = x.shape[-2] // 2 - H // 2
    c0 = x.shape[-1] // 2 - W // 2
    return x[..., r0:r0 + H, c0:c0 + W]


def psnr(target, recon):
    mse = np.mean((np.asarray(target, np.float64) - np.asarray(recon, np.float64)) ** 2)
    return 10.0 * np.log10(1.0 / mse)


def ssim_channel(x, y, win=7):
    """scikit-image SSIM of two 2-D images (data_range 1, uniform window)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    npix = win * win
    cov = npix / (npix - 1.0)
    f = lambda a: ndimage.uniform_filter(a, size=win)
    ux, uy = f(x), f(y)
    vx = cov * (f(x * x) - ux * ux)
    vy = cov * (f(y * y) - uy * uy)
    vxy = cov * (f(x * y) - ux * uy)
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    s = ((2 * ux * uy + c1) * (2 * vxy + c2)) / ((ux * ux + uy * uy + c1) * (vx + vy + c2))
    p = (win - 1) // 2
    return s[p:-p, p:-p].mean()


def ssim_channel_bruteforce(x, y, win=7):
    """Window-by-window restatement of ssim_channel (interior positions only)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    p = (win - 1) // 2
    npix = win * win
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    vals = []
    for r in range(p, x.shape[0] - p):
        for c in range(p, x.shape[1] - p):
            a = x[r - p:r + p + 1, c - p:c + p + 1].ravel()
            b = y[r - p:r + p + 1, c - p:c + p + 1].ravel()
            ua, ub = a.mean(), b.mean()
            va = ((a - ua) ** 2).sum() / (npix - 1)
            vb = ((b - ub) ** 2).sum() / (npix - 1)
            vab = ((a - ua) * (b - ub)).sum() / (npix - 1)
            vals.append(((2 * ua * ub + c1) * (2 * vab + c2)) /
                        ((ua * ua + ub * ub + c1) * (va + vb + c2)))
    return float(np.mean(vals))


def evaluate(recons, targets, H, W):
    """Per-plane PSNR and SSIM (train_2d_gaussian_rgb.py:98-112): recons [D, C, rH, rW]
    are clamped to [0, 1] and centre-cropped to H x W; targets [D, C, H, W]."""
    out = {"psnr": [], "ssim": []}
    for d in range(min(len(recons), len(targets))):
        r = crop_center(np.clip(np.asarray(recons[d], np.float64), 0.0, 1.0), H, W)
        t = np.asarray(targets[d], np.float64)
        out["psnr"].append(psnr(t, r))
        out["ssim"].append(float(np.mean([ssim_channel(t[c], r[c]) for c in range(t.shape[0])])))
    return out
