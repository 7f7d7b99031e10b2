"""Fused pad -> glue -> BLASM -> crop on the gfx950 kernels of libcvqpg.

``HologramOptics`` replaces, for one training step, the chain
soft_window_pad (model_2d_gaussian_rgb.py:186-241) -> reconstruct glue
(train_2d_gaussian_rgb.py:89-95) -> propagator.reconstruct (utils/propagator.py:224-290)
-> crop_center(size=(H, W)) (train_2d_gaussian_rgb.py:193): it maps the raster field
[C, H, W] to the per-plane intensities [D, C, H, W] that the loss consumes, and its
adjoint maps dL/dI back to dL/d(field).  ``propagate`` is the differentiable entry
point (a torch.autograd.Function around the two C calls).
"""
import torch

from . import _lib


class HologramOptics:
    def __init__(self, img_size, pad_size, wavelengths, pixel_pitch, distances,
                 aperture_size=None, grid="odak", aperture=None):
        """``aperture`` (optional): a custom Fourier aperture [2 pH, 2 pW] on the centred
        grid (the propagator's zero-padded ``set_aperture`` tensor), used by the full-plane
        ``reconstruct_planes`` in place of the circular ``aperture_size``."""
        self.W, self.H = int(img_size[0]), int(img_size[1])
        self.pH, self.pW = int(pad_size[0]), int(pad_size[1])
        self.wavelengths = [float(w) for w in wavelengths]
        self.C = len(self.wavelengths)
        self.distances = [float(z) for z in distances]
        self.D = len(self.distances)
        self.pitch = float(pixel_pitch)
        self.aperture = float(aperture_size) if aperture_size is not None else float(max(self.pH, self.pW))
        self.grid = {"odak": 0, "fft": 1}[grid]
        self.custom_aperture = None
        if aperture is not None:
            a = torch.as_tensor(aperture, dtype=torch.float32).contiguous()
            if tuple(a.shape) != (2 * self.pH, 2 * self.pW):
                raise ValueError(f"aperture must be {2 * self.pH}x{2 * self.pW}, got "
                                 f"{tuple(a.shape)}")
            self.custom_aperture = a
        self._ws = {}

    def _workspace(self, device, full_plane=0):
        key = (str(device), full_plane)
        if key not in self._ws:
            nb = _lib.lib().cvqpg_blasm_workspace_size(self.C, self.D, self.H, self.W, self.pH,
                                                       self.pW, full_plane)
            self._ws[key] = torch.empty(int(nb), dtype=torch.uint8, device=device)
        return self._ws[key]

    def _common(self):
        return (self.C, self.H, self.W, self.pH, self.pW, self.D, _lib.doubles(self.wavelengths),
                _lib.doubles(self.distances), self.pitch, self.aperture, self.grid)

    def forward(self, field, keep_field=True):
        """field: complex64 [C, H, W] on the GPU -> (intensity [D,C,H,W], plane fields or None)."""
        field = field.resolve_conj().contiguous()
        assert field.shape == (self.C, self.H, self.W) and field.dtype == torch.complex64
        dev = field.device
        ws = self._workspace(dev)
        inten = torch.empty(self.D, self.C, self.H, self.W, device=dev)
        out = torch.empty(self.D, self.C, self.H, self.W, dtype=torch.complex64, device=dev) \
            if keep_field else None
        _lib.check(_lib.lib().cvqpg_blasm_forward(
            _lib.ptr(torch.view_as_real(field)), *self._common(), _lib.ptr(ws), ws.numel(),
            None if out is None else _lib.ptr(torch.view_as_real(out)), _lib.ptr(inten),
            _lib.stream_handle(dev)))
        return inten, out

    def backward(self, field, grad_intensity, out_field):
        field = field.resolve_conj().contiguous()
        g = grad_intensity.contiguous().float()
        dev = field.device
        ws = self._workspace(dev)
        grad = torch.empty(self.C, self.H, self.W, dtype=torch.complex64, device=dev)
        _lib.check(_lib.lib().cvqpg_blasm_backward(
            _lib.ptr(torch.view_as_real(field)), _lib.ptr(g), _lib.ptr(torch.view_as_real(out_field)),
            *self._common(), _lib.ptr(ws), ws.numel(), _lib.ptr(torch.view_as_real(grad)),
            _lib.stream_handle(dev)))
        return grad

    def _plane_common(self):
        return (self.C, self.pH, self.pW, self.D, _lib.doubles(self.wavelengths),
                _lib.doubles(self.distances), self.pitch, self.aperture, self.grid)

    def _aperture_on(self, device):
        """The custom aperture on ``device`` (cached), or None."""
        if self.custom_aperture is None:
            return None
        if self.custom_aperture.device != torch.device(device):
            self.custom_aperture = self.custom_aperture.to(device)
        return self.custom_aperture

    def _aperture_common(self):
        return (self.C, self.pH, self.pW, self.D, _lib.doubles(self.wavelengths),
                _lib.doubles(self.distances), self.pitch, self.grid)

    def reconstruct_planes(self, phase, amplitude, keep_field=False):
        """propagator.reconstruct for one frame per channel over the full padded plane:
        phase, amplitude [C, pH, pW] -> intensity [D, C, pH, pW], and with ``keep_field``
        also the complex plane fields [D, C, pH, pW] the backward needs (else None)."""
        dev = phase.device
        ws = self._workspace(dev, 1)
        inten = torch.empty(self.D, self.C, self.pH, self.pW, device=dev)
        out = torch.empty(self.D, self.C, self.pH, self.pW, dtype=torch.complex64, device=dev) \
            if keep_field else None
        ph, am = phase.contiguous().float(), amplitude.contiguous().float()
        of = None if out is None else _lib.ptr(torch.view_as_real(out))
        ap = self._aperture_on(dev)
        if ap is None:
            _lib.check(_lib.lib().cvqpg_blasm_reconstruct(
                _lib.ptr(ph), _lib.ptr(am), *self._plane_common(), _lib.ptr(ws), ws.numel(), of,
                _lib.ptr(inten), _lib.stream_handle(dev)))
        else:
            _lib.check(_lib.lib().cvqpg_blasm_reconstruct_aperture(
                _lib.ptr(ph), _lib.ptr(am), _lib.ptr(ap), *self._aperture_common(), _lib.ptr(ws),
                ws.numel(), of, _lib.ptr(inten), _lib.stream_handle(dev)))
        return (inten, out) if keep_field else inten

    def reconstruct_planes_backward(self, phase, amplitude, grad_intensity, out_field):
        """Adjoint of ``reconstruct_planes``: (dL/dphase, dL/damplitude), each [C, pH, pW]."""
        dev = phase.device
        ws = self._workspace(dev, 1)
        gp = torch.empty(self.C, self.pH, self.pW, device=dev)
        ga = torch.empty_like(gp)
        ph, am = phase.contiguous().float(), amplitude.contiguous().float()
        gi = grad_intensity.contiguous().float()
        of = _lib.ptr(torch.view_as_real(out_field))
        ap = self._aperture_on(dev)
        if ap is None:
            _lib.check(_lib.lib().cvqpg_blasm_reconstruct_backward(
                _lib.ptr(ph), _lib.ptr(am), _lib.ptr(gi), of, *self._plane_common(),
                _lib.ptr(ws), ws.numel(), _lib.ptr(gp), _lib.ptr(ga), _lib.stream_handle(dev)))
        else:
            _lib.check(_lib.lib().cvqpg_blasm_reconstruct_aperture_backward(
                _lib.ptr(ph), _lib.ptr(am), _lib.ptr(ap), _lib.ptr(gi), of,
                *self._aperture_common(), _lib.ptr(ws), ws.numel(), _lib.ptr(gp), _lib.ptr(ga),
                _lib.stream_handle(dev)))
        return gp, ga


class _ReconstructPlanes(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phase, amplitude, optics):
        phase = phase.contiguous().float()
        amplitude = amplitude.contiguous().float()
        inten, out = optics.reconstruct_planes(phase, amplitude, keep_field=True)
        ctx.save_for_backward(phase, amplitude, out)
        ctx.optics = optics
        return inten

    @staticmethod
    def backward(ctx, grad_inten):
        phase, amplitude, out = ctx.saved_tensors
        gp, ga = ctx.optics.reconstruct_planes_backward(phase, amplitude, grad_inten, out)
        return gp, ga, None


def reconstruct_planes(phase, amplitude, optics):
    """Differentiable phase, amplitude [C, pH, pW] -> intensities [D, C, pH, pW] over the
    whole padded plane (one frame per channel)."""
    return _ReconstructPlanes.apply(phase, amplitude, optics)


class _Propagate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, field, optics):
        inten, out = optics.forward(field, keep_field=True)
        ctx.save_for_backward(field, out)
        ctx.optics = optics
        return inten

    @staticmethod
    def backward(ctx, grad_inten):
        field, out = ctx.saved_tensors
        return ctx.optics.backward(field, grad_inten, out), None


def propagate(field, optics):
    """Differentiable field [C,H,W] complex64 -> intensities [D,C,H,W]."""
    return _Propagate.apply(field, optics)
