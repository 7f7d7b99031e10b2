"""One full training step on the CPU in torch autograd (FP64 by default).

Restates the step order of train_2d_gaussian_rgb.py:187-211 (and the gray twin):
prologue (model:83-153, 424-439) -> binned raster (cuda/forward.cu:62-162) ->
soft_window_pad (model:186-241) -> glue (train:89-95) -> BLASM to D planes
(utils/propagator.py:248-290) -> centre crop + clamp -> per-plane 3-term L2 +
SSIM (train:191-197) -> backward -> nan_to_num + clip (train:205-209) -> Adan
(utils/optimizer.py:105-262) -> means cosine schedule (train:58-70).

The raster gradients come from torch autograd of the differentiable binned
forward below, which is independent of the closed-form adjoint in
``oracle.raster.render_backward``; the two are compared in the tests.
"""
import torch

from . import adan, binning, loss, model, optics, raster


def render_binned_diff(ri, width, height, bins):
    """Differentiable binned forward over the (pixel, Gaussian) pairs of ``bins``."""
    pxi, pyi, gi = raster._pairs(bins, width, height)
    means, inv_cov = ri["means"], ri["inv_cov"]
    colours, phase, opac, curv, area = ri["colours"], ri["phase"], ri["opac"], ri["curv"], ri["area"]
    k, dxy2 = ri["k"], ri["dxy2"]
    C = colours.shape[1]
    rdt = means.dtype
    cdt = torch.complex128 if rdt == torch.float64 else torch.complex64
    dx = means[gi, 0] - pxi.to(rdt)
    dy = means[gi, 1] - pyi.to(rdt)
    _, _, _, alpha, base = raster.pair_terms(dx, dy, inv_cov[gi, 0], inv_cov[gi, 1],
                                             inv_cov[gi, 2], opac[gi], curv[gi], area[gi], dxy2)
    pid = pyi * width + pxi
    chans = []
    for c in range(C):
        th = phase[gi, c] + k[c].to(rdt) * base
        amp = colours[gi, c] * alpha
        contrib = torch.complex(amp * torch.cos(th), amp * torch.sin(th))
        chans.append(torch.zeros(height * width, dtype=cdt).index_add(0, pid, contrib))
    return torch.stack(chans).view(C, height, width)


class StepConfig:
    def __init__(self, img_size, wavelengths, pixel_pitch, pad_size, distances,
                 aperture_size=None, grid="odak", curv_mode="scale_aware", grayscale=None):
        self.img_size = tuple(img_size)
        self.wavelengths = list(wavelengths)
        self.pixel_pitch = float(pixel_pitch)
        self.pad_size = tuple(pad_size)
        self.distances = list(distances)
        self.aperture_size = aperture_size
        self.grid = grid
        self.curv_mode = curv_mode
        self.grayscale = (len(self.wavelengths) == 1) if grayscale is None else grayscale


def forward(raw, cfg, prop=None, dtype=torch.float64):
    """Returns (intensities [D,C,pH,pW], field [C,H,W], bins, raster inputs)."""
    W, H = cfg.img_size
    ri = model.raster_inputs(raw, cfg.img_size, cfg.wavelengths, cfg.pixel_pitch,
                             cfg.curv_mode, grayscale=cfg.grayscale)
    bins = binning.bin_gaussians(ri["means"].detach().to(torch.float32).numpy(),
                                 ri["radii"].numpy(), W, H)
    field = render_binned_diff(ri, W, H, bins)
    padded = model.soft_window_pad(field, cfg.pad_size)
    u, _, _ = model.glue(padded)
    if prop is None:
        prop = optics.BLASM(cfg.pad_size, cfg.wavelengths, cfg.pixel_pitch, cfg.distances,
                            cfg.aperture_size, cfg.grid)
    return prop.reconstruct(u), field, bins, ri


def loss_and_grads(raw, cfg, targets, masks, noises=None, prop=None, frozen=()):
    """Forward + backward.  ``raw``: dict of FP64 leaf tensors.  Returns
    (loss, grads dict (None for frozen), intensities)."""
    W, H = cfg.img_size
    leaves = {n: raw[n].detach().clone().requires_grad_(n not in frozen) for n in raw}
    intens, _, _, _ = forward(leaves, cfg, prop)
    L = loss.step_loss(intens, targets, masks, H, W, noises)
    L.backward()
    grads = {n: (None if n in frozen else leaves[n].grad) for n in raw}
    return L.detach(), grads, intens.detach()


def train_step(raw, cfg, targets, masks, opt_state, noises=None, prop=None, frozen=()):
    """Full step; updates ``raw`` in place and returns (loss, grad_norm)."""
    L, grads, _ = loss_and_grads(raw, cfg, targets, masks, noises, prop, frozen)
    grads, norm = adan.clip_grads(grads)
    opt_state.step(raw, grads)
    return float(L), norm
