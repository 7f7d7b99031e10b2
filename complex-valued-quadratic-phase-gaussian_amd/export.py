"""Phase-only hologram export (DPAC, 8-bit) and per-plane evaluation on the GPU.

Mirrors the numeric core of the reference's export script and evaluation:
  * ``dpac`` and the 8-bit encoding                  export_poh.py:110-119, 173-178
  * reconstruction of the 8-bit POH through the Fourier aperture of radius
    min(img_size), and of the complex field as in training   export_poh.py:185-190
  * PSNR / SSIM per plane (scikit-image defaults)   train_2d_gaussian_rgb.py:98-112
Image files, LPIPS / FLIP / CVVDP and the command line are not part of this module.
Every numeric step runs in libcvqpg (``cvqpg_soft_window_pad``, ``cvqpg_dpac_encode``,
``cvqpg_blasm_reconstruct``, ``cvqpg_plane_metrics``); nothing falls back to the CPU.
"""
import math

import torch

from . import _lib
from .optics import HologramOptics

TWO_PI = 2.0 * math.pi


def soft_window_pad(field, pad_size):
    """Guard-band padding of a complex64 field [C, H, W] -> [C, pH, pW]
    (model_2d_gaussian_rgb.py:186-241), no gradient."""
    field = field.detach().resolve_conj().contiguous()
    assert field.dtype == torch.complex64 and field.dim() == 3
    C, H, W = field.shape
    pH, pW = int(pad_size[0]), int(pad_size[1])
    out = torch.empty(C, pH, pW, dtype=torch.complex64, device=field.device)
    _lib.check(_lib.lib().cvqpg_soft_window_pad(
        _lib.ptr(torch.view_as_real(field)), C, H, W, pH, pW, _lib.ptr(torch.view_as_real(out)),
        _lib.stream_handle(field.device)))
    return out


def dpac_levels(field):
    """8-bit DPAC levels uint8 [C, pH, pW] of a padded complex64 field (export_poh.py:
    110-119, 173-178: phase mod 2 pi, amplitude clamped to [0, 1], phase -+ arccos(A) on
    the checkerboard, round(. / 2 pi * 256) mod 256)."""
    field = field.detach().resolve_conj().contiguous()
    C, pH, pW = field.shape
    levels = torch.empty(C, pH, pW, dtype=torch.uint8, device=field.device)
    _lib.check(_lib.lib().cvqpg_dpac_encode(
        _lib.ptr(torch.view_as_real(field)), C, pH, pW, _lib.ptr(levels),
        _lib.stream_handle(field.device)))
    return levels


def plane_metrics(recons, targets, H, W):
    """Per-plane PSNR and SSIM on the device (no host synchronisation): recons
    [D, C, rH, rW] (clamped to [0, 1] and centre-cropped to H x W), targets [D, C, H, W].
    Returns two device tensors [D]."""
    recons = recons.detach().contiguous().float()
    targets = targets.detach().contiguous().float()
    D, C, rH, rW = recons.shape
    assert targets.shape == (D, C, H, W)
    L = _lib.lib()
    ws = torch.empty(int(L.cvqpg_metrics_workspace_size(D, C, H, W)), dtype=torch.uint8,
                     device=recons.device)
    psnr = torch.empty(D, device=recons.device)
    ssim = torch.empty(D, device=recons.device)
    _lib.check(L.cvqpg_plane_metrics(_lib.ptr(recons), rH, rW, _lib.ptr(targets), D, C, H, W,
                                     _lib.ptr(ws), ws.numel(), _lib.ptr(psnr), _lib.ptr(ssim),
                                     _lib.stream_handle(recons.device)))
    return psnr, ssim


def evaluate(recons, targets, H, W):
    """train_2d_gaussian_rgb.py:98-112 restricted to PSNR and SSIM: lists per plane."""
    psnr, ssim = plane_metrics(recons, targets, H, W)
    return {"psnr": [float(v) for v in psnr.cpu()], "ssim": [float(v) for v in ssim.cpu()]}


def make_optics(args_prop, img_size, aperture_size):
    """propagator of export_poh.py:95-107 (one frame per channel, BLASM)."""
    return HologramOptics(img_size, args_prop.pad_size, args_prop.wavelengths,
                          args_prop.pixel_pitch, args_prop.distances,
                          aperture_size=aperture_size)


def export_poh(scene, args_prop, img_size, aperture_size=None, targets=None):
    """The numeric core of export_poh.py:main for a fitted scene: render, pad, encode,
    reconstruct the 8-bit POH and the complex field, and (with targets [D, C, H, W])
    evaluate both.  Returns a dict with ``levels`` uint8 [C, pH, pW], ``recon_poh`` and
    ``recon_complex`` [D, C, pH, pW] and, with targets, ``poh`` / ``complex`` metrics."""
    W, H = int(img_size[0]), int(img_size[1])
    aperture = aperture_size if aperture_size is not None else min(W, H)
    with torch.no_grad():
        field = soft_window_pad(scene.render_unpadded(), args_prop.pad_size)
        levels = dpac_levels(field)
        poh_phase = levels.float() * (TWO_PI / 256.0)
        recon_poh = make_optics(args_prop, img_size, aperture).reconstruct_planes(
            poh_phase, torch.ones_like(poh_phase))
        phase = torch.remainder(torch.angle(field), TWO_PI)
        amplitude = torch.clamp(torch.abs(field), 0.0, 1.0)
        recon_complex = make_optics(args_prop, img_size, args_prop.aperture_size).reconstruct_planes(
            phase, amplitude)
    out = {"levels": levels, "recon_poh": recon_poh, "recon_complex": recon_complex,
           "poh_aperture_size": aperture,
           "encoding": "DPAC, 8-bit, gray level g = phase 2*pi*g/256"}
    if targets is not None:
        out["complex"] = evaluate(recon_complex, targets, H, W)
        out["poh"] = evaluate(recon_poh, targets, H, W)
    return out
