"""Gaussian scene model with the reference's interface (model_2d_gaussian_rgb.py).

* ``Gaussians2D``      -- seven raw parameter tensors, activations, covariance and
                          inverse (model_2d_gaussian_rgb.py:10-183; the grayscale twin
                          differs only in C = 1 and the curvature bound's wavelength,
                          model_2d_gaussian_grayscale.py:57,69,96).
* ``soft_window_pad``  -- guard-band padding with a Tukey taper (:186-241).
* ``Scene2D``          -- render() -> padded complex field [C, pH, pW] (:244-444), with
                          the ``renderer`` switch of :261-267: ``'hip'`` (alias
                          ``'cuda'``, the value the reference's --renderer flag uses)
                          runs the gfx950 rasterizer of this package; ``'torch'`` is the
                          reference's PyTorch tile renderer (3 sigma, 64 px tiles),
                          selectable only explicitly -- it is never used as a fallback.
* ``make_trainable_2d`` (:447-465).

Checkpoints keep the reference's ``.pth`` keys (:155-180).
"""
import math

import numpy as np
import torch
import torch.nn.functional as F

from .rasterizer import qpf_render

TWO_PI = 2.0 * math.pi
PARAM_NAMES = ("means_2d", "pre_act_scales", "pre_act_rotation", "colours", "pre_act_phase",
               "pre_act_opacities", "curvature")


class Gaussians2D(torch.nn.Module):
    def __init__(self, num_points, img_size, device, args_prop, channels=None, init=True):
        super().__init__()
        self.device = device
        self.img_size = tuple(img_size)
        self.args_prop = args_prop
        C = channels if channels is not None else len(args_prop.wavelengths)
        self.channels = C
        self.grayscale = C == 1
        data = self._initialize_2d(num_points, self.img_size, C) if init else {
            n: torch.zeros(s) for n, s in self._shapes(num_points, C).items()}
        for n in PARAM_NAMES:
            self.register_parameter(n, torch.nn.Parameter(data[n], requires_grad=False))
        self.curv_nyquist_frac = getattr(args_prop, "curv_nyquist", 0.9)
        self.curv_mode = "scale_aware"
        self.curv_const_max = 200.0
        self.to(device)

    @staticmethod
    def _shapes(n, C):
        return dict(means_2d=(n, 2), pre_act_scales=(n, 2), pre_act_rotation=(n,),
                    colours=(n, C), pre_act_phase=(n, C), pre_act_opacities=(n,), curvature=(n,))

    @staticmethod
    def _initialize_2d(n, img_size, C):
        """Random initialisation of model_2d_gaussian_rgb.py:41-77."""
        W, H = img_size
        m = torch.rand((n, 2))
        norm = torch.stack([m[:, 0] * 2 - 1, m[:, 1] * 2 - 1], 1).clamp(-0.999, 0.999)
        return dict(means_2d=torch.atanh(norm),
                    colours=torch.rand((n, C)),
                    pre_act_scales=torch.log(torch.rand((n, 2)) * 3.5 + 1.5),
                    pre_act_rotation=torch.rand((n,)) * 2 * np.pi - np.pi,
                    pre_act_phase=torch.zeros((n, C)),
                    pre_act_opacities=torch.full((n,), -0.5),
                    curvature=torch.zeros((n,)))

    def curvature_wavelength(self):
        wl = self.args_prop.wavelengths
        return float(wl[0]) if self.grayscale else float(min(wl))

    def apply_activations(self):
        """model_2d_gaussian_rgb.py:83-112."""
        W, H = self.img_size
        mt = torch.tanh(self.means_2d)
        means = torch.stack([(mt[:, 0] + 1) * 0.5 * W, (mt[:, 1] + 1) * 0.5 * H], dim=1)
        scales = torch.exp(self.pre_act_scales) + 0.1
        rotation = self.pre_act_rotation
        phase = self.pre_act_phase % TWO_PI
        opac = torch.sigmoid(self.pre_act_opacities)
        dxy = float(self.args_prop.pixel_pitch)
        sigma = scales.max(dim=1).values
        if self.curv_mode == "off":
            curv = torch.zeros_like(self.curvature)
        elif self.curv_mode == "constant":
            curv = self.curv_const_max * torch.tanh(self.curvature)
        elif self.curv_mode == "scale_aware":
            cmax = self.curv_nyquist_frac * self.curvature_wavelength() / (6.0 * sigma * dxy ** 2 + 1e-20)
            curv = cmax * torch.tanh(self.curvature)
        else:
            raise ValueError(f"Unknown curv_mode: {self.curv_mode}")
        return scales, rotation, phase, opac, means, curv

    @staticmethod
    def invert_cov_2D(c00, c01, c11):
        """model_2d_gaussian_rgb.py:114-125."""
        det = (c00 * c11 - c01 * c01).clamp(min=1e-10)
        inv_det = 1.0 / det
        return c11 * inv_det, -c01 * inv_det, c00 * inv_det

    @staticmethod
    def compute_2d_covariance_elements(scales, rotation):
        """model_2d_gaussian_rgb.py:127-153 (without the cache the reference invalidates
        on every render anyway, :443)."""
        c, s = torch.cos(rotation), torch.sin(rotation)
        sx2, sy2 = scales[:, 0] ** 2, scales[:, 1] ** 2
        return (sx2 * c * c + sy2 * s * s + 0.1, (sx2 - sy2) * c * s,
                sx2 * s * s + sy2 * c * c + 0.1)

    def params(self):
        return [getattr(self, n) for n in PARAM_NAMES]

    def state(self):
        d = {n: getattr(self, n).detach().cpu() for n in PARAM_NAMES}
        d["img_size"] = self.img_size
        return d

    def save_gaussians(self, path):
        """Same dict keys as model_2d_gaussian_rgb.py:155-167."""
        torch.save(self.state(), path)

    def load_gaussians(self, path):
        """model_2d_gaussian_rgb.py:169-180 (tensors only: weights_only=True)."""
        sd = torch.load(path, map_location="cpu", weights_only=True)
        with torch.no_grad():
            for n in PARAM_NAMES:
                getattr(self, n).copy_(sd[n].to(self.device))

    def __len__(self):
        return self.means_2d.shape[0]


def pad_offsets(H, W, pH, pW):
    pad_h = (pH - H) // 2
    pad_w = (pW - W) // 2
    return pad_h, pH - H - pad_h, pad_w, pW - W - pad_w


def taper_1d(total, inner_start, inner_size, device, dtype=torch.float32):
    """Tukey ramp of model_2d_gaussian_rgb.py:224-235."""
    w = torch.zeros(total, device=device, dtype=dtype)
    w[inner_start:inner_start + inner_size] = 1.0
    if inner_start > 0:
        t = torch.arange(inner_start, device=device, dtype=dtype)
        w[:inner_start] = 0.5 * (1.0 - torch.cos(math.pi * (t + 1.0) / inner_start))
    rs = inner_start + inner_size
    nr = total - rs
    if nr > 0:
        t = torch.arange(nr, device=device, dtype=dtype)
        w[rs:] = 0.5 * (1.0 + torch.cos(math.pi * (t + 1.0) / nr))
    return w


def soft_window_pad(hologram, pad_size):
    """model_2d_gaussian_rgb.py:186-241 (PyTorch; the fused HIP step applies the same
    padding inside its first FFT pass)."""
    C, H, W = hologram.shape
    pH, pW = int(pad_size[0]), int(pad_size[1])
    ph, phb, pw, pwr = pad_offsets(H, W, pH, pW)
    with torch.no_grad():
        rep = torch.complex(F.pad(hologram.real.unsqueeze(0), (pw, pwr, ph, phb), mode="replicate")[0],
                            F.pad(hologram.imag.unsqueeze(0), (pw, pwr, ph, phb), mode="replicate")[0])
    zero = torch.complex(F.pad(hologram.real, (pw, pwr, ph, phb)), F.pad(hologram.imag, (pw, pwr, ph, phb)))
    mask = torch.zeros(1, pH, pW, device=hologram.device)
    mask[:, ph:ph + H, pw:pw + W] = 1.0
    padded = zero + rep * (1.0 - mask)
    wy = taper_1d(pH, ph, H, hologram.device)
    wx = taper_1d(pW, pw, W, hologram.device)
    return padded * torch.outer(wy, wx).unsqueeze(0)


class Scene2D:
    def __init__(self, gaussians, args_prop):
        self.gaussians = gaussians
        self.args_prop = args_prop
        self.device = gaussians.device
        self.wavelengths = torch.tensor(args_prop.wavelengths, dtype=torch.float32, device=self.device)
        self.tile_size = getattr(args_prop, "tile_size", 64)
        self.gauss_batch = getattr(args_prop, "gauss_batch", 1024)
        self.phase_iso = getattr(args_prop, "phase_iso", False)
        self.renderer = getattr(args_prop, "renderer", "hip")
        if self.renderer in ("hip", "cuda"):
            if self.phase_iso:
                raise ValueError("--phase_iso is only supported by the torch renderer")
            if not str(self.device).startswith("cuda"):
                raise ValueError("--renderer hip needs a GPU device")
        elif self.renderer != "torch":
            raise ValueError(f"unknown renderer {self.renderer!r}")

    def raster_inputs(self):
        """Prologue of model_2d_gaussian_rgb.py:383-386 and :428-436."""
        g = self.gaussians
        scales, rotation, phase, opac, means, curv = g.apply_activations()
        c00, c01, c11 = g.compute_2d_covariance_elements(scales, rotation)
        i00, i01, i11 = g.invert_cov_2D(c00, c01, c11)
        with torch.no_grad():
            mid = 0.5 * (c00 + c11)
            disc = torch.sqrt(((c00 - c11) * 0.5) ** 2 + c01 ** 2)
            radii = (5.0 * torch.sqrt(mid + disc)).ceil().to(torch.int32)
        return dict(means=means, inv_cov=torch.stack([i00, i01, i11], -1), colours=g.colours,
                    phase=phase, opac=opac, curv=curv,
                    area=(scales[:, 0] * scales[:, 1]).abs(), radii=radii, scales=scales,
                    cov=(c00, c01, c11))

    def render_unpadded(self):
        W, H = self.gaussians.img_size
        k = 2 * np.pi / self.wavelengths
        dxy2 = float(self.args_prop.pixel_pitch) ** 2
        ri = self.raster_inputs()
        if self.renderer in ("hip", "cuda"):
            return qpf_render(ri["means"], ri["inv_cov"], ri["colours"], ri["phase"], ri["opac"],
                              ri["curv"], ri["area"], ri["radii"], k.contiguous(), dxy2, W, H)
        return self._render_torch(ri, k, dxy2, W, H)

    def _render_torch(self, ri, k, dxy2, W, H):
        """The reference's PyTorch tile loop (model_2d_gaussian_rgb.py:393-419); with
        ``args_prop.grad_ckpt`` each tile's body runs under activation checkpointing
        (:407-416), so autograd keeps one tile's output instead of its batches."""
        from torch.utils.checkpoint import checkpoint
        C = len(self.wavelengths)
        c00, c01, c11 = ri["cov"]
        cf = (ri["colours"] * torch.exp(1j * ri["phase"])).to(torch.complex64)
        radius = 3.0 * torch.sqrt(0.5 * (c00 + c11) + torch.sqrt(((c00 - c11) * 0.5) ** 2 + c01 ** 2))
        ts = self.tile_size
        diag = math.sqrt(2 * (ts / 2) ** 2)
        holo = torch.zeros((C, H, W), dtype=torch.complex64, device=self.device)
        scales = ri["scales"]
        area = (scales[:, 0] * scales[:, 1]).abs()
        tensors = (ri["means"], ri["inv_cov"], ri["opac"], ri["curv"], area, cf, radius)
        ckpt = bool(getattr(self.args_prop, "grad_ckpt", False)) and torch.is_grad_enabled()
        for y0 in range(0, H, ts):
            y1 = min(y0 + ts, H)
            for x0 in range(0, W, ts):
                x1 = min(x0 + ts, W)
                if ckpt:
                    acc = checkpoint(self._tile_torch, (x0, y0, x1, y1), k, dxy2, diag, *tensors,
                                     use_reentrant=False)
                else:
                    acc = self._tile_torch((x0, y0, x1, y1), k, dxy2, diag, *tensors)
                if acc is not None:
                    holo[:, y0:y1, x0:x1] = holo[:, y0:y1, x0:x1] + acc.view(C, y1 - y0, x1 - x0)
        return holo

    def _tile_torch(self, box, k, dxy2, diag, means, inv, opac, curv, area, cf, radius):
        """Complex contribution [C, pixels] of the tile ``box`` = (x0, y0, x1, y1): the
        Gaussians whose 3-sigma disc reaches the tile's circumscribed circle, evaluated
        ``gauss_batch`` at a time; None when no Gaussian reaches the tile."""
        x0, y0, x1, y1 = box
        dev = self.device
        C = len(self.wavelengths)
        cx, cy = (x0 + x1) / 2.0, (y0 + y1) / 2.0
        mask = (means[:, 0] - cx) ** 2 + (means[:, 1] - cy) ** 2 < (radius + diag) ** 2
        if not bool(mask.any()):
            return None
        ys = torch.arange(y0, y1, device=dev, dtype=torch.float32)
        xs = torch.arange(x0, x1, device=dev, dtype=torch.float32)
        gy, gx = torch.meshgrid(ys, xs, indexing="ij")
        pts = torch.stack([gx, gy], -1).reshape(1, -1, 2)
        idx = torch.nonzero(mask).squeeze(1)
        acc = torch.zeros((C, pts.shape[1]), dtype=torch.complex64, device=dev)
        for g0 in range(0, idx.numel(), self.gauss_batch):
            b = idx[g0:g0 + self.gauss_batch]
            diff = pts - means[b].unsqueeze(1)
            dx, dy = diff[..., 0], diff[..., 1]
            m = dx * dx * inv[b, 0, None] + 2 * dx * dy * inv[b, 1, None] + dy * dy * inv[b, 2, None]
            alphas = opac[b, None] * torch.exp(torch.clamp(-0.5 * m, min=-50))
            if self.phase_iso:
                base = 0.5 * curv[b, None] * (dx * dx + dy * dy) * dxy2
            else:
                base = 0.5 * curv[b, None] * m * area[b, None] * dxy2
            mod = torch.polar(alphas.unsqueeze(0).expand(C, -1, -1), k.view(C, 1, 1) * base.unsqueeze(0))
            acc = acc + torch.einsum("cbp,bc->cp", mod, cf[b])
        return acc

    def render(self, img_size=None):
        """Padded complex hologram [C, pH, pW] (model_2d_gaussian_rgb.py:441-444)."""
        return soft_window_pad(self.render_unpadded(), self.args_prop.pad_size)


def make_trainable_2d(gaussians):
    """model_2d_gaussian_rgb.py:447-465 (the cache-invalidation hooks are not needed:
    nothing is cached between renders)."""
    for p in gaussians.params():
        p.requires_grad_()
