"""Band-limited angular spectrum (BLASM) propagation, restated in torch FP64.

Restates
  * ``propagator._reconstruct_impl``     utils/propagator.py:248-290
  * ``custom``                          utils/propagator.py:41-64
  * ``circular_binary_mask``            utils/propagator.py:10-38
  * odak 0.2.7/0.2.8 helpers it calls (not vendored, not installable here; the
    restatement follows odak's published source as recalled, so these
    conventions are "parity unpinned" against odak itself):
      - ``zero_pad``      centred 2x zero padding, start = n - n // 2
      - ``crop_center``   centre crop; size None -> start n // 4, length n // 2;
                          explicit size (h, w) -> [n//2 - h//2, n//2 + h//2)
      - ``get_band_limited_angular_spectrum_kernel``: frequency grid
        ``linspace(-1/(2dx) + 0.5/(2X), 1/(2dx) - 0.5/(2X), n)`` with X = dx * n
        (the "odak" grid; ``grid="fft"`` selects exact FFT bins (k - n/2) / X),
        ``FY, FX = meshgrid(fx, fy, indexing='ij')``, phase
        2 pi z sqrt(1/lambda^2 - FX^2 - FY^2) evaluated in FP64 (odak does it in
        FP32, which costs ~2e-3 rad rms at these sizes), band limit
        |FX| < 1/(lambda sqrt((2 z / X)^2 + 1)) and the same for FY with Y.
        Evanescent points (negative radicand) are set to zero.

Pinned by (tests/test_oracle_optics.py): z = 0 gives the identity on the
cropped field; |H| = 1 inside the band so the padded-grid energy is conserved
(Parseval); H(z1) H(z2) = H(z1 + z2) for the exact-bin grid; one FFT-bin plane
wave is multiplied by H at that bin; the odak grid's end points and spacing;
numpy.fft (pocketfft) as the FFT.
"""
import math

import numpy as np
import torch


def freq_grid(n, dx, grid="odak"):
    """FP64 frequencies along one axis, in fftshift (centred) order."""
    X = dx * float(n)
    if grid == "odak":
        return np.linspace(-1.0 / (2 * dx) + 0.5 / (2 * X), 1.0 / (2 * dx) - 0.5 / (2 * X), n)
    if grid == "fft":
        return (np.arange(n) - n // 2) / X
    raise ValueError(grid)


def transfer_function(nu, nv, dx, wavelength, distance, grid="odak", phase_dtype=np.float64):
    """Centred BLASM transfer function [nu, nv] complex128 (odak convention).
    ``phase_dtype=np.float32`` evaluates the frequency grid and the phase
    2 pi z sqrt(1/lambda^2 - FX^2 - FY^2) in FP32, as odak does on float32 tensors (the
    reference's precision, not the oracle's)."""
    fx = freq_grid(nu, dx, grid).astype(phase_dtype)   # rows (dim 0) -> named FY by odak
    fy = freq_grid(nv, dx, grid).astype(phase_dtype)   # cols (dim 1) -> named FX by odak
    FY, FX = np.meshgrid(fx, fy, indexing="ij")
    one = phase_dtype(1.0)
    rad = one / phase_dtype(wavelength) ** 2 - (FX ** 2 + FY ** 2)
    ok = rad >= 0
    phase = (phase_dtype(2 * np.pi) * np.sqrt(np.where(ok, rad, phase_dtype(0.0)))
             * phase_dtype(distance)).astype(np.float64)
    x = dx * float(nu)
    y = dx * float(nv)
    fx_max = 1.0 / math.sqrt((2 * distance * (1.0 / x)) ** 2 + 1) / wavelength
    fy_max = 1.0 / math.sqrt((2 * distance * (1.0 / y)) ** 2 + 1) / wavelength
    filt = (np.abs(FX) < fx_max) & (np.abs(FY) < fy_max) & ok
    return np.where(filt, np.exp(1j * phase), 0.0)


def circular_mask(px, py, r):
    """utils/propagator.py:10-38 (note: linspace(-n/2, n/2, n), not integer steps)."""
    x = np.linspace(-px / 2.0, px / 2.0, px)
    y = np.linspace(-py / 2.0, py / 2.0, py)
    X, Y = np.meshgrid(x, y, indexing="ij")
    return (np.sqrt(X ** 2 + Y ** 2) < r).astype(np.float64)


def zero_pad2(u):
    """odak zero_pad to 2x, centred. u: [..., n, m]."""
    n, m = u.shape[-2:]
    out = torch.zeros(u.shape[:-2] + (2 * n, 2 * m), dtype=u.dtype)
    s0, s1 = n - n // 2, m - m // 2
    out[..., s0:s0 + n, s1:s1 + m] = u
    return out


def crop_center(u, size=None):
    """odak crop_center (see module docstring for the index convention)."""
    n, m = u.shape[-2:]
    if size is None:
        return u[..., n // 4:n // 4 + n // 2, m // 4:m // 4 + m // 2]
    h, w = size
    return u[..., n // 2 - h // 2:n // 2 + h // 2, m // 2 - w // 2:m // 2 + w // 2]


class BLASM:
    """FP64 propagator for the padded field [C, pH, pW] to D planes.

    ``reconstruct(u)`` returns intensities [D, C, pH, pW] (the frame sum with the
    identity laser power of train:378-388), differentiable through torch.
    """

    def __init__(self, pad_size, wavelengths, pixel_pitch, distances, aperture_size=None,
                 grid="odak", phase_dtype=np.float64, aperture=None):
        """``aperture`` (optional, [pH, pW]): a custom Fourier aperture at the hologram's
        resolution, zero-padded to the 2x grid as set_aperture does
        (utils/propagator.py:199-200); otherwise the circular mask of ``aperture_size``."""
        self.pH, self.pW = pad_size
        self.wavelengths = list(wavelengths)
        self.dx = float(pixel_pitch)
        self.distances = [float(z) for z in distances]
        nu, nv = 2 * self.pH, 2 * self.pW
        if aperture is not None:
            self.aperture = zero_pad2(torch.as_tensor(aperture, dtype=torch.float64))
        else:
            if aperture_size is None:
                aperture_size = max(self.pH, self.pW)
            self.aperture = torch.as_tensor(circular_mask(nu, nv, aperture_size))
        self.H = torch.stack([torch.stack([
            torch.as_tensor(transfer_function(nu, nv, self.dx, wl, z, grid, phase_dtype))
            for wl in self.wavelengths]) for z in self.distances])           # [D, C, nu, nv]

    def field_at(self, u, d, c):
        """Complex field at plane d for channel c: custom() + crop_center (propagator.py:284-288)."""
        U = zero_pad2(u[c])
        ap = self.aperture.to(U.dtype)
        S = torch.fft.fftshift(torch.fft.fft2(U)) * ap
        S = (self.H[d, c] * ap) * S
        out = torch.fft.ifft2(torch.fft.ifftshift(S))
        return crop_center(out)

    def reconstruct(self, u):
        D, C = len(self.distances), u.shape[0]
        return torch.stack([torch.stack([self.field_at(u, d, c).abs() ** 2 for c in range(C)])
                            for d in range(D)])
