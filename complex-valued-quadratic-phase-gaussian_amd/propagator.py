"""Band-limited angular spectrum propagation with the reference's interface
(utils/propagator.py:67-290).

``propagator(resolution, wavelengths, pixel_pitch, number_of_frames, ..., distances,
aperture_size, device)`` and ``reconstruct(hologram_phases, amplitude, no_grad)`` ->
intensities [frames, depths, channels, pH, pW] keep the reference's signature and
meaning.  The transfer function H = exp(i 2 pi z sqrt(1/lambda^2 - fx^2 - fy^2)) is
evaluated with an FP64 phase (the reference's odak kernel evaluates it in FP32,
which is off by ~2e-3 rad rms at these sizes).  The frequency grid is odak's
half-sample ``linspace(-1/(2dx) + 1/(4X), 1/(2dx) - 1/(4X), n)`` by default
(``grid='odak'``) or the exact FFT bins (``grid='fft'``).

``reconstruct`` runs on the gfx950 BLASM kernels of libcvqpg when ``backend='hip'``
(the default; differentiable, with the adjoint kernels behind a torch.autograd.Function,
so ``no_grad=False`` trains through it as train_2d_gaussian_rgb.py:94 does) and on
torch.fft with a cached transfer function when ``backend='torch'`` (the reference's own
structure, kept as the PyTorch baseline).  There is no fallback between the two: the
HIP backend raises if the library is missing.  A custom aperture tensor
(utils/propagator.py:195-200, zero-padded to the 2x grid) runs on both backends (on HIP
through cvqpg_blasm_reconstruct_aperture, which applies it, as custom() does, to the
spectrum and to the transfer function).
"""
import math

import torch

from . import optics as optics_mod


def freq_grid(n, dx, grid="odak", device="cpu"):
    X = dx * float(n)
    if grid == "odak":
        return torch.linspace(-1.0 / (2 * dx) + 0.5 / (2 * X), 1.0 / (2 * dx) - 0.5 / (2 * X), n,
                              dtype=torch.float64, device=device)
    if grid == "fft":
        return (torch.arange(n, dtype=torch.float64, device=device) - n // 2) / X
    raise ValueError(grid)


def transfer_function(nu, nv, dx, wavelength, distance, grid="odak", device="cpu"):
    """Centred BLASM transfer function [nu, nv] complex64 with an FP64 phase
    (odak's get_propagation_kernel for 'Bandlimited Angular Spectrum')."""
    fx = freq_grid(nu, dx, grid, device)
    fy = freq_grid(nv, dx, grid, device)
    FY, FX = torch.meshgrid(fx, fy, indexing="ij")
    rad = 1.0 / wavelength ** 2 - (FX ** 2 + FY ** 2)
    ok = rad >= 0
    ph = 2 * math.pi * torch.sqrt(torch.clamp(rad, min=0.0)) * float(distance)
    ph = torch.remainder(ph, 2 * math.pi)
    fx_max = 1.0 / math.sqrt((2 * distance / (dx * nu)) ** 2 + 1) / wavelength
    fy_max = 1.0 / math.sqrt((2 * distance / (dx * nv)) ** 2 + 1) / wavelength
    filt = (FX.abs() < fx_max) & (FY.abs() < fy_max) & ok
    H = torch.polar(filt.to(torch.float64), ph)
    return H.to(torch.complex64)


def circular_binary_mask(px, py, r, cx=0, cy=0, device="cpu"):
    """utils/propagator.py:10-38."""
    x = torch.linspace(-px / 2.0, px / 2.0, px, device=device)
    y = torch.linspace(-py / 2.0, py / 2.0, py, device=device)
    X, Y = torch.meshgrid(x, y, indexing="ij")
    return (((X - cx) ** 2 + (Y - cy) ** 2) ** 0.5 < r).float()


class propagator:  # noqa: N801  (reference class name)
    def __init__(self, resolution=(1920, 1080), wavelengths=(515e-9,), pixel_pitch=8e-6,
                 number_of_frames=1, number_of_depth_layers=1, volume_depth=1e-2,
                 image_location_offset=5e-3, propagation_type="Bandlimited Angular Spectrum",
                 laser_channel_power=None, aperture=None, aperture_size=None, distances=None,
                 aperture_samples=(20, 20, 5, 5), device=torch.device("cpu"), grid="odak",
                 backend="hip"):
        if propagation_type != "Bandlimited Angular Spectrum":
            raise ValueError("only 'Bandlimited Angular Spectrum' propagation is implemented")
        self.device = torch.device(device)
        self.pixel_pitch = float(pixel_pitch)
        self.wavelengths = [float(w) for w in wavelengths]
        self.resolution = [int(resolution[0]), int(resolution[1])]
        self.number_of_frames = number_of_frames
        self.number_of_channels = len(self.wavelengths)
        self.grid = grid
        self.backend = backend
        if distances is None:
            self.distances = torch.linspace(-volume_depth / 2.0, volume_depth / 2.0,
                                            number_of_depth_layers) + image_location_offset
        else:
            self.distances = torch.as_tensor(distances)
        self.number_of_depth_layers = int(self.distances.shape[0])
        self.channel_power = (torch.eye(number_of_frames, self.number_of_channels)
                              if laser_channel_power is None else torch.as_tensor(laser_channel_power))
        self.channel_power = self.channel_power.float().cpu()
        nu, nv = 2 * self.resolution[0], 2 * self.resolution[1]
        self.custom_aperture = aperture is not None
        if aperture is None:
            if aperture_size is None:
                aperture_size = max(self.resolution)
            self.aperture_size = float(aperture_size)
            self.aperture = circular_binary_mask(nu, nv, aperture_size, device=self.device)
        else:
            # utils/propagator.py:199-200: the aperture at the hologram's resolution,
            # zero-padded (odak zero_pad, centred) to the 2x grid
            a = torch.as_tensor(aperture, dtype=torch.float32, device=self.device)
            pH, pW = self.resolution
            if a.shape[-2:] != (pH, pW):
                raise ValueError(f"aperture must be {pH}x{pW}, got {tuple(a.shape[-2:])}")
            self.aperture = torch.zeros(nu, nv, device=self.device)
            self.aperture[pH - pH // 2:pH - pH // 2 + pH, pW - pW // 2:pW - pW // 2 + pW] = \
                a.reshape(pH, pW)
            self.aperture_size = None
        self.aperture_is_open = bool(self.aperture.min() >= 1.0)
        if backend not in ("hip", "torch"):
            raise ValueError(f"backend must be 'hip' or 'torch', got {backend!r}")
        self._kernels = None
        self._optics = None

    def distance_list(self):
        return [float(z) for z in self.distances.double()]

    def kernels(self):
        """[D, C, 2pH, 2pW] transfer functions in FFT (unshifted) order, with the aperture."""
        if self._kernels is None:
            nu, nv = 2 * self.resolution[0], 2 * self.resolution[1]
            ks = []
            for z in self.distance_list():
                row = []
                for wl in self.wavelengths:
                    H = transfer_function(nu, nv, self.pixel_pitch, wl, z, self.grid, self.device)
                    if not self.aperture_is_open:
                        # custom() applies the aperture to the spectrum and to H
                        # (utils/propagator.py:59-61): a factor aperture^2
                        H = H * (self.aperture * self.aperture)
                    row.append(torch.fft.ifftshift(H))
                ks.append(torch.stack(row))
            self._kernels = torch.stack(ks)
        return self._kernels

    def lit(self):
        """(frame, channel) pairs with non-zero laser power (utils/propagator.py:280-283)."""
        return [(f, c, float(self.channel_power[f, c])) for f in range(self.number_of_frames)
                for c in range(self.number_of_channels) if float(self.channel_power[f, c]) != 0]

    def reconstruct(self, hologram_phases, amplitude=None, no_grad=True):
        """utils/propagator.py:224-246: hologram phases [F, pH, pW] (a leading batch dim
        of 1 is squeezed) and amplitudes [C, pH, pW] -> intensities [F, D, C, pH, pW]."""
        impl = self._reconstruct_hip if self.backend == "hip" else self._reconstruct_torch
        if no_grad:
            with torch.no_grad():
                return impl(hologram_phases, amplitude)
        return impl(hologram_phases, amplitude)

    def optics(self):
        """The HIP propagation plan of this propagator (full padded planes)."""
        if self._optics is None:
            pH, pW = self.resolution
            self._optics = optics_mod.HologramOptics(
                (pW, pH), (pH, pW), self.wavelengths, self.pixel_pitch, self.distance_list(),
                aperture_size=self.aperture_size, grid=self.grid,
                aperture=self.aperture if self.custom_aperture else None)
        return self._optics

    def _reconstruct_hip(self, phases, amplitude):
        """utils/propagator.py:248-290 on the gfx950 kernels: one differentiable
        cvqpg_blasm_reconstruct per frame, channel c lit with power[f, c] (unlit channels
        contribute zero, as the reference's skip leaves them)."""
        if phases.dim() > 3:
            phases = phases.squeeze(0)
        pH, pW = self.resolution
        D, C, F = self.number_of_depth_layers, self.number_of_channels, self.number_of_frames
        if amplitude is None:
            amplitude = torch.ones(C, pH, pW, device=phases.device)
        opt = self.optics()
        power = self.channel_power
        eye = power.shape == (C, C) and bool(torch.equal(power, torch.eye(C)))
        if eye:      # frame c lights channel c only: one call for every frame
            inten = optics_mod.reconstruct_planes(phases[:C], amplitude, opt)   # [D, C, pH, pW]
            zero = torch.zeros_like(inten[:, 0])
            return torch.stack([torch.stack([inten[:, c] if c == f else zero for c in range(C)],
                                            dim=1) for f in range(F)])
        frames = []
        for f in range(F):
            pw = power[f].to(phases.device, amplitude.dtype)
            if bool((pw == 0).all()):
                frames.append(torch.zeros(D, C, pH, pW, device=phases.device))
                continue
            ph = phases[f].unsqueeze(0).expand(C, pH, pW)
            frames.append(optics_mod.reconstruct_planes(ph, amplitude * pw.view(C, 1, 1), opt))
        return torch.stack(frames)

    def _reconstruct_torch(self, phases, amplitude):
        """utils/propagator.py:248-290 on torch.fft; the aperture is folded into H."""
        if phases.dim() > 3:
            phases = phases.squeeze(0)
        pH, pW = self.resolution
        D, C = self.number_of_depth_layers, self.number_of_channels
        if amplitude is None:
            amplitude = torch.ones(C, pH, pW, device=phases.device)
        Hk = self.kernels()
        out = []
        for f in range(self.number_of_frames):
            per_d = [[None] * C for _ in range(D)]
            for (ff, c, pw) in self.lit():
                if ff != f:
                    continue
                a = pw * amplitude[c]
                u = torch.polar(a, phases[f])
                U = torch.zeros(2 * pH, 2 * pW, dtype=torch.complex64, device=u.device)
                U[pH - pH // 2:pH - pH // 2 + pH, pW - pW // 2:pW - pW // 2 + pW] = u
                S = torch.fft.fft2(U)
                for d in range(D):
                    o = torch.fft.ifft2(S * Hk[d, c])
                    o = o[pH // 2:pH // 2 + pH, pW // 2:pW // 2 + pW]
                    per_d[d][c] = o.real ** 2 + o.imag ** 2
            zero = torch.zeros(pH, pW, device=phases.device)
            out.append(torch.stack([torch.stack([x if x is not None else zero for x in row])
                                    for row in per_d]))
        return torch.stack(out)
