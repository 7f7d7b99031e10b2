"""Drop-in replacement for the reference's cuda/rasterizer.py on MI355X.

``qpf_render`` keeps the signature, argument meaning, return type and error
behaviour of cuda/rasterizer.py:62-71 (``QPFRender`` of :36-59): inputs are the
activated per-Gaussian tensors, the result is the complex64 field [C, H, W], and
the backward returns gradients for the first seven arguments and ``None`` for
radii, k, dxy2, width and height.  The work is done by libcvqpg (hand-written
gfx950 HIP kernels) through its C ABI; there is no CPU or PyTorch fallback.

Like cuda/raster.cu:71 the eager forward reads the key count P back to the host
once to size its key buffers; the graph-captured training step uses
``cvqpg_qpf_render_forward`` with a pre-sized capacity instead.
"""
import math

import torch
from torch.autograd import Function

from . import _lib

TILE = 16
_capacity_hint = {}


def _check_input(t, name, n):
    """cuda/raster.cu:22-28: device tensor, contiguous, float32/int32, first dim n."""
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be CUDA")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    if t.dtype not in (torch.float32, torch.int32):
        raise RuntimeError(f"{name} must be float32/int32")
    if t.size(0) != n:
        raise RuntimeError(f"{name} has {t.size(0)} rows, expected {n}")


def workspace_bytes(n, channels, width, height, capacity):
    return int(_lib.lib().cvqpg_raster_workspace_size(n, channels, width, height, capacity))


class Binning:
    """Key list of one forward pass (sorted ids, tile ranges, workspace)."""

    def __init__(self, ws, capacity, point_list, ranges, num_pairs):
        self.ws, self.capacity = ws, capacity
        self.point_list, self.ranges, self.num_pairs = point_list, ranges, num_pairs


def bin_gaussians(means_2d, radii, width, height, channels=3):
    """Count, emit, sort and range the (tile, Gaussian) keys on the GPU."""
    L = _lib.lib()
    n = means_2d.shape[0]
    dev = means_2d.device
    key = (n, width, height)
    cap = max(1024, int(_capacity_hint.get(key, 16 * max(n, 1)) * 1.125))
    stream = _lib.stream_handle(dev)
    num = torch.empty(1, dtype=torch.int32, device=dev)
    ws = torch.empty(workspace_bytes(n, channels, width, height, cap), dtype=torch.uint8, device=dev)
    _lib.check(L.cvqpg_bin_count(_lib.ptr(means_2d), _lib.ptr(radii), n, width, height,
                                 _lib.ptr(ws), ws.numel(), cap, _lib.ptr(num), stream))
    P = int(num.item())  # host read-back of P, as cuda/raster.cu:71
    if P > cap:
        cap = P
        ws = torch.empty(workspace_bytes(n, channels, width, height, cap), dtype=torch.uint8,
                         device=dev)
        _lib.check(L.cvqpg_bin_count(_lib.ptr(means_2d), _lib.ptr(radii), n, width, height,
                                     _lib.ptr(ws), ws.numel(), cap, _lib.ptr(num), stream))
    _capacity_hint[key] = P
    T = math.ceil(width / TILE) * math.ceil(height / TILE)
    point_list = torch.empty(max(P, 1), dtype=torch.int32, device=dev)
    ranges = torch.empty(T, 2, dtype=torch.int32, device=dev)
    _lib.check(L.cvqpg_bin_sort(_lib.ptr(means_2d), _lib.ptr(radii), n, width, height,
                                _lib.ptr(ws), ws.numel(), cap, _lib.ptr(point_list),
                                _lib.ptr(ranges), stream))
    return Binning(ws, cap, point_list, ranges, P)


def render_forward(means_2d, inv_cov, colours, phase, opacities, curvature, area, k, dxy2,
                   width, height, bins):
    L = _lib.lib()
    n, C = colours.shape
    field = torch.empty(C, height, width, 2, dtype=torch.float32, device=means_2d.device)
    st = _lib.stream_handle(means_2d.device)
    ws, nb, cap = _lib.ptr(bins.ws), bins.ws.numel(), bins.capacity
    _lib.check(L.cvqpg_raster_pack(
        _lib.ptr(means_2d), _lib.ptr(inv_cov), _lib.ptr(colours), _lib.ptr(phase),
        _lib.ptr(opacities), _lib.ptr(curvature), _lib.ptr(area), _lib.ptr(k), n, C, float(dxy2),
        width, height, ws, nb, cap, st))
    timer.wrap("k_render_fwd", lambda: _lib.check(L.cvqpg_render_forward_packed(
        n, C, width, height, ws, nb, cap, _lib.ptr(bins.point_list), _lib.ptr(bins.ranges),
        _lib.ptr(field), st)))
    return torch.view_as_complex(field)


class StageTimer:
    """Optional HIP-event timing of single raster kernels on the launch stream.

    When ``enabled``, the adjoint tile kernel (k_render_bwd) and the forward tile
    kernel (k_render_fwd) are bracketed by events on the stream they are launched on;
    ``mean_ms(name)`` returns their average duration."""

    def __init__(self):
        self.enabled = False
        self.events = {}

    def wrap(self, name, fn):
        if not self.enabled:
            return fn()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        out = fn()
        b.record()
        self.events.setdefault(name, []).append((a, b))
        return out

    def mean_ms(self, name):
        ev = self.events.get(name, [])
        if not ev:
            return None
        torch.cuda.synchronize()
        return sum(a.elapsed_time(b) for a, b in ev) / len(ev)

    def reset(self):
        self.events = {}


timer = StageTimer()


def render_backward(grad_field, means_2d, inv_cov, colours, phase, opacities, curvature, area, k,
                    dxy2, width, height, bins):
    L = _lib.lib()
    n, C = colours.shape
    dev = means_2d.device
    g = torch.view_as_real(grad_field.resolve_conj().contiguous()).contiguous()
    outs = [torch.empty(n, 2, device=dev), torch.empty(n, 3, device=dev),
            torch.empty(n, C, device=dev), torch.empty(n, C, device=dev),
            torch.empty(n, device=dev), torch.empty(n, device=dev), torch.empty(n, device=dev)]
    if n == 0:
        return [o.zero_() for o in outs]
    st = _lib.stream_handle(dev)
    ws, nb, cap = _lib.ptr(bins.ws), bins.ws.numel(), bins.capacity
    _lib.check(L.cvqpg_raster_pack(
        _lib.ptr(means_2d), _lib.ptr(inv_cov), _lib.ptr(colours), _lib.ptr(phase),
        _lib.ptr(opacities), _lib.ptr(curvature), _lib.ptr(area), _lib.ptr(k), n, C, float(dxy2),
        width, height, ws, nb, cap, st))
    timer.wrap("k_render_bwd", lambda: _lib.check(L.cvqpg_render_backward_tiles(
        _lib.ptr(g), _lib.ptr(k), n, C, width, height, ws, nb, cap, _lib.ptr(bins.point_list),
        _lib.ptr(bins.ranges), st)))
    _lib.check(L.cvqpg_render_backward_reduce(
        _lib.ptr(colours), _lib.ptr(curvature), _lib.ptr(area), n, C, float(dxy2), width, height,
        ws, nb, cap, *[_lib.ptr(o) for o in outs], st))
    return outs


class QPFRender(Function):
    """Autograd wrapper with the signature of cuda/rasterizer.py:36-59."""

    @staticmethod
    def forward(ctx, means_2d, inv_cov, colours, phase, opacities, curvature, area,
                radii, k, dxy2, width, height):
        C = colours.shape[1]
        if C not in (1, 3):
            raise RuntimeError(f"qpf rasterizer supports 1 or 3 channels, got {C}")
        args = [t.contiguous() for t in (means_2d, inv_cov, colours, phase, opacities,
                                         curvature, area)]
        n = args[0].shape[0]
        radii_i = radii.contiguous().int()
        for t, name in zip(args + [radii_i], ("means_2d", "inv_cov", "colours", "phase",
                                                "opacities", "curvature", "area", "radii")):
            _check_input(t, name, n)
        kk = k.contiguous().float()
        if kk.numel() != C:
            raise RuntimeError("k must have num_channels entries")
        bins = bin_gaussians(args[0], radii_i, int(width), int(height), C)
        field = render_forward(*args, kk, dxy2, int(width), int(height), bins)
        ctx.save_for_backward(*args, kk, bins.point_list, bins.ranges, bins.ws)
        ctx.cfg = (float(dxy2), int(width), int(height), bins.capacity, bins.num_pairs)
        return field

    @staticmethod
    def backward(ctx, grad_out):
        (means, inv_cov, colours, phase, opac, curv, area, k, point_list, ranges,
         ws) = ctx.saved_tensors
        dxy2, W, H, cap, P = ctx.cfg
        bins = Binning(ws, cap, point_list, ranges, P)
        d = render_backward(grad_out, means, inv_cov, colours, phase, opac, curv, area, k, dxy2,
                            W, H, bins)
        return d[0], d[1], d[2], d[3], d[4], d[5], d[6], None, None, None, None, None


def qpf_render(means_2d, inv_cov, colours, phase, opacities, curvature, area,
               radii, k, dxy2, width, height):
    """Complex field [C, H, W] of N Gaussians (linear coherent sum, quadratic phase).

    means_2d [N,2] pixel coords, inv_cov [N,3] = (inv00, inv01, inv11), colours/phase [N,C],
    opacities/curvature/area [N] (area = |sx*sy|), radii [N] culling radius in px,
    k [C] wavenumbers, dxy2 = pixel_pitch**2.
    """
    return QPFRender.apply(means_2d, inv_cov, colours, phase, opacities, curvature, area,
                           radii, k, dxy2, width, height)


def qpf_render_reference(means_2d, inv_cov, colours, phase, opacities, curvature, area,
                         k, dxy2, width, height, pixels_per_chunk=1 << 22):
    """Dense evaluation of every (Gaussian, pixel) term in PyTorch, without culling: the
    testing helper the reference exports next to qpf_render (cuda/rasterizer.py:74-90).
    Differentiable; works a band of image rows at a time so that at most
    ``pixels_per_chunk`` (Gaussian, pixel) terms are live.  Never used by qpf_render."""
    n = means_2d.shape[0]
    i00, i01, i11 = inv_cov.unbind(1)
    xs = torch.arange(width, device=means_2d.device, dtype=means_2d.dtype)
    ex = means_2d[:, 0, None, None] - xs                    # [N, 1, W]
    scale = (0.5 * dxy2) * curvature * area                 # per Gaussian, times the Mahalanobis
    rows = max(1, pixels_per_chunk // max(1, n * width))
    bands = []
    for y0 in range(0, height, rows):
        ys = torch.arange(y0, min(height, y0 + rows), device=means_2d.device, dtype=means_2d.dtype)
        ey = means_2d[:, 1, None, None] - ys[:, None]       # [N, R, 1]
        q = (i00[:, None, None] * ex * ex + 2.0 * i01[:, None, None] * ex * ey
             + i11[:, None, None] * ey * ey)               # [N, R, W]
        weight = opacities[:, None, None] * torch.exp(torch.clamp(-0.5 * q, min=-50.0))
        chirp = scale[:, None, None] * q
        per_channel = []
        for c in range(colours.shape[1]):
            theta = phase[:, c, None, None] + k[c] * chirp
            amp = colours[:, c, None, None] * weight
            per_channel.append(torch.complex((amp * torch.cos(theta)).sum(0),
                                             (amp * torch.sin(theta)).sum(0)))
        bands.append(torch.stack(per_channel))              # [C, R, W]
    return torch.cat(bands, dim=1)
