"""Gaussian prologue, guard-band padding, reconstruct glue and the torch tile
renderer, restated in torch (dtype-generic; the oracle runs them in FP64).

Restates
  * ``Gaussians2D.apply_activations``        model_2d_gaussian_rgb.py:83-112
                                             (gray twin: model_2d_gaussian_grayscale.py:96)
  * ``compute_2d_covariance_elements``       model_2d_gaussian_rgb.py:127-153
  * ``invert_cov_2D``                        model_2d_gaussian_rgb.py:114-125
  * ``Scene2D._render_cuda`` prologue        model_2d_gaussian_rgb.py:424-439
  * ``soft_window_pad``                      model_2d_gaussian_rgb.py:186-241
  * ``reconstruct`` glue                     train_2d_gaussian_rgb.py:89-95
  * torch tile renderer (3 sigma, 64 px)     model_2d_gaussian_rgb.py:269-422

Pinned by (tests/test_oracle_model.py): closed forms (rotation 0 gives a
diagonal covariance; inverse times covariance is the identity; area = sx*sy;
radii of an isotropic Gaussian = ceil(5 sqrt(s^2 + 0.1))), the Tukey window's
end values and symmetry, the centre of the padded field equal to the input,
the replicate band equal to the nearest edge value, the glue's identity for
|h| <= 1 and unit-modulus projection for |h| > 1, and the tile renderer
against the dense sum (3-sigma culling error bound).
"""
import math

import numpy as np
import torch
import torch.nn.functional as F

TWO_PI = 2.0 * math.pi


def apply_activations(raw, img_size, wavelengths, pixel_pitch, curv_mode="scale_aware",
                      curv_nyquist=0.9, curv_const_max=200.0, grayscale=False):
    """raw: dict of pre-activation tensors (means_2d, pre_act_scales, pre_act_rotation,
    colours, pre_act_phase, pre_act_opacities, curvature).  Returns
    (scales, rotation, phase, opacities, means_2d, curvature) as in :83-112."""
    W, H = img_size
    mt = torch.tanh(raw["means_2d"])
    means = torch.stack([(mt[:, 0] + 1) * 0.5 * W, (mt[:, 1] + 1) * 0.5 * H], dim=1)
    scales = torch.exp(raw["pre_act_scales"]) + 0.1
    rotation = raw["pre_act_rotation"]
    phase = raw["pre_act_phase"] % TWO_PI
    opac = torch.sigmoid(raw["pre_act_opacities"])
    wl = float(wavelengths[0]) if grayscale else float(min(wavelengths))
    dxy = float(pixel_pitch)
    sigma = scales.max(dim=1).values
    if curv_mode == "off":
        curv = torch.zeros_like(raw["curvature"])
    elif curv_mode == "constant":
        curv = curv_const_max * torch.tanh(raw["curvature"])
    elif curv_mode == "scale_aware":
        cmax = curv_nyquist * wl / (6.0 * sigma * (dxy ** 2) + 1e-20)
        curv = cmax * torch.tanh(raw["curvature"])
    else:
        raise ValueError(f"Unknown curv_mode: {curv_mode}")
    return scales, rotation, phase, opac, means, curv


def covariance(scales, rotation):
    c, s = torch.cos(rotation), torch.sin(rotation)
    sx, sy = scales[:, 0], scales[:, 1]
    sx2, sy2 = sx * sx, sy * sy
    c2, s2, cs = c * c, s * s, c * s
    return sx2 * c2 + sy2 * s2 + 0.1, (sx2 - sy2) * cs, sx2 * s2 + sy2 * c2 + 0.1


def invert_cov(c00, c01, c11):
    det = (c00 * c11 - c01 * c01).clamp(min=1e-10)
    inv_det = 1.0 / det
    return c11 * inv_det, -c01 * inv_det, c00 * inv_det


def raster_inputs(raw, img_size, wavelengths, pixel_pitch, curv_mode="scale_aware",
                  grayscale=False):
    """Everything ``qpf_render`` receives (model:384-386 + :424-439).
    Returns dict(means, inv_cov, colours, phase, opac, curv, area, radii, k, dxy2)."""
    scales, rotation, phase, opac, means, curv = apply_activations(
        raw, img_size, wavelengths, pixel_pitch, curv_mode, grayscale=grayscale)
    c00, c01, c11 = covariance(scales, rotation)
    i00, i01, i11 = invert_cov(c00, c01, c11)
    inv_cov = torch.stack([i00, i01, i11], dim=-1)
    area = (scales[:, 0] * scales[:, 1]).abs()
    with torch.no_grad():
        mid = 0.5 * (c00 + c11)
        disc = torch.sqrt(((c00 - c11) * 0.5) ** 2 + c01 ** 2)
        radii = (5.0 * torch.sqrt(mid + disc)).ceil().to(torch.int32)
    wl = torch.tensor(list(wavelengths), dtype=means.dtype)
    k = 2 * np.pi / wl
    return dict(means=means, inv_cov=inv_cov, colours=raw["colours"], phase=phase, opac=opac,
                curv=curv, area=area, radii=radii, k=k, dxy2=float(pixel_pitch) ** 2,
                scales=scales, cov=(c00, c01, c11))


def taper_1d(total, inner_start, inner_size, dtype=torch.float64):
    """Tukey ramp of soft_window_pad (model:224-235)."""
    w = torch.zeros(total, dtype=dtype)
    w[inner_start:inner_start + inner_size] = 1.0
    if inner_start > 0:
        t = torch.arange(inner_start, dtype=dtype)
        w[:inner_start] = 0.5 * (1.0 - torch.cos(math.pi * (t + 1.0) / inner_start))
    right_start = inner_start + inner_size
    n_right = total - right_start
    if n_right > 0:
        t = torch.arange(n_right, dtype=dtype)
        w[right_start:] = 0.5 * (1.0 + torch.cos(math.pi * (t + 1.0) / n_right))
    return w


def pad_offsets(H, W, pH, pW):
    pad_h = (pH - H) // 2
    pad_w = (pW - W) // 2
    return pad_h, pH - H - pad_h, pad_w, pW - W - pad_w


def soft_window_pad(h, pad_size):
    """model:186-241: zero pad with gradient in the centre, replicate pad (no grad) in the
    guard band, times the outer product of two Tukey ramps."""
    C, H, W = h.shape
    pH, pW = pad_size
    pad_h, pad_hb, pad_w, pad_wr = pad_offsets(H, W, pH, pW)
    rdt = h.real.dtype
    with torch.no_grad():
        rep = torch.complex(
            F.pad(h.real.unsqueeze(0), (pad_w, pad_wr, pad_h, pad_hb), mode="replicate").squeeze(0),
            F.pad(h.imag.unsqueeze(0), (pad_w, pad_wr, pad_h, pad_hb), mode="replicate").squeeze(0))
    zero = torch.complex(F.pad(h.real, (pad_w, pad_wr, pad_h, pad_hb)),
                         F.pad(h.imag, (pad_w, pad_wr, pad_h, pad_hb)))
    mask = torch.zeros(1, pH, pW, dtype=rdt)
    mask[:, pad_h:pad_h + H, pad_w:pad_w + W] = 1.0
    padded = zero + rep * (1.0 - mask)
    wy = taper_1d(pH, pad_h, H, rdt)
    wx = taper_1d(pW, pad_w, W, rdt)
    return padded * torch.outer(wy, wx).unsqueeze(0)


def glue(h):
    """train:89-95 + odak generate_complex_field: u = clamp(|h|,0,1) e^{i (angle(h) mod 2pi)}.
    Returns (u, phase_map, amplitude)."""
    ph = torch.angle(h) % TWO_PI
    amp = torch.clamp(torch.abs(h), min=0.0, max=1.0)
    u = torch.complex(amp * torch.cos(ph), amp * torch.sin(ph))
    return u, ph, amp


def render_tile_torch(ri, width, height, tile_size=64, gauss_batch=1024):
    """The reference's default ``--renderer torch`` path (model:374-419) for the
    Mahalanobis QPF, restated: 3-sigma circle culling per 64-px tile, batches of
    ``gauss_batch`` Gaussians, polar + einsum.  Differentiable; dtype follows
    the inputs.  Returns complex [C, H, W] (unpadded)."""
    means, opac, curv = ri["means"], ri["opac"], ri["curv"]
    c00, c01, c11 = ri["cov"]
    i00, i01, i11 = ri["inv_cov"][:, 0], ri["inv_cov"][:, 1], ri["inv_cov"][:, 2]
    k, dxy2, scales = ri["k"], ri["dxy2"], ri["scales"]
    colours, phase = ri["colours"], ri["phase"]
    C = colours.shape[1]
    rdt = means.dtype
    cdt = torch.complex128 if rdt == torch.float64 else torch.complex64
    cf = (colours * torch.exp(1j * phase)).to(cdt)
    mid = 0.5 * (c00 + c11)
    disc = torch.sqrt(((c00 - c11) * 0.5) ** 2 + c01 ** 2)
    radius = 3.0 * torch.sqrt(mid + disc)
    diag = math.sqrt(2 * (tile_size / 2) ** 2)
    holo = torch.zeros((C, height, width), dtype=cdt)
    area_all = (scales[:, 0] * scales[:, 1]).abs()
    for y0 in range(0, height, tile_size):
        y1 = min(y0 + tile_size, height)
        for x0 in range(0, width, tile_size):
            x1 = min(x0 + tile_size, width)
            cx, cy = (x0 + x1) / 2.0, (y0 + y1) / 2.0
            mask = (means[:, 0] - cx) ** 2 + (means[:, 1] - cy) ** 2 < (radius + diag) ** 2
            ys = torch.arange(y0, y1, dtype=rdt)
            xs = torch.arange(x0, x1, dtype=rdt)
            gy, gx = torch.meshgrid(ys, xs, indexing="ij")
            pts = torch.stack([gx, gy], -1).reshape(1, -1, 2)
            acc = torch.zeros((C, pts.shape[1]), dtype=cdt)
            if not bool(mask.any()):
                continue
            idx = torch.nonzero(mask).squeeze(1)
            for g0 in range(0, idx.numel(), gauss_batch):
                b = idx[g0:g0 + gauss_batch]
                diff = pts - means[b].unsqueeze(1)
                dx, dy = diff[..., 0], diff[..., 1]
                mahal = dx * dx * i00[b, None] + 2 * dx * dy * i01[b, None] + dy * dy * i11[b, None]
                alphas = opac[b, None] * torch.exp(torch.clamp(-0.5 * mahal, min=-50))
                base = 0.5 * curv[b, None] * mahal * area_all[b, None] * dxy2
                mod = torch.polar(alphas.unsqueeze(0).expand(C, -1, -1),
                                  k.to(rdt).view(C, 1, 1) * base.unsqueeze(0))
                acc = acc + torch.einsum("cbp,bc->cp", mod, cf[b])
            holo[:, y0:y1, x0:x1] = holo[:, y0:y1, x0:x1] + acc.view(C, y1 - y0, x1 - x0)
    return holo
