"""Quadratic-phase Gaussian raster, forward and adjoint, restated in torch FP64 (CPU).

Per pixel p = (x, y) (integer coordinates) and Gaussian g:
  d      = mu_g - p
  mahal  = i00 dx^2 + 2 i01 dx dy + i11 dy^2
  G      = exp(max(-mahal / 2, -50))
  alpha  = opacity_g * G
  base   = 0.5 * curvature_g * mahal * area_g * dxy2
  U[c,p] = sum_g colour[g,c] * alpha * exp(i (phase[g,c] + k_c * base))

Restates
  * ``qpf_render_reference``  cuda/rasterizer.py:74-90   (dense, no culling)
  * ``renderTileKernel``      cuda/forward.cu:62-162      (binned: only the pairs
                              whose tile lies in the Gaussian's 5-sigma rectangle)
  * ``renderBackwardKernel``  cuda/backward.cu:11-168     (closed-form adjoint)

Pinned by (tests/test_oracle_raster.py):
  * closed forms: one isotropic flat Gaussian is a * o * e^{i phi} *
    exp(-r^2 / (2 s^2)); with curvature the phase is the chirp
    phi + k * 0.5 * c * mahal * area * dxy2;
  * linearity / superposition in the Gaussian set;
  * the adjoint against FP64 central finite differences of the forward and
    against torch autograd of the dense forward (independent of the closed-form
    expressions), and the dot-product test <J v, w> = <v, J^T w>.
"""
import numpy as np
import torch

from . import binning

F64 = torch.float64


def _t(a, dtype=F64):
    if isinstance(a, torch.Tensor):
        return a.detach().to(dtype=dtype, device="cpu")
    return torch.as_tensor(np.asarray(a), dtype=dtype)


def pair_terms(dx, dy, i00, i01, i11, opac, curv, area, dxy2):
    """Shared per-pair quantities (mahal, G, alpha, base, clamp-active mask)."""
    mahal = dx * (i00 * dx + i01 * dy) + dy * (i01 * dx + i11 * dy)
    raw = -0.5 * mahal
    G = torch.exp(torch.clamp(raw, min=-50.0))
    alpha = opac * G
    base = 0.5 * curv * mahal * area * dxy2
    return mahal, raw, G, alpha, base


def render_dense(means, inv_cov, colours, phase, opac, curv, area, k, dxy2, width, height,
                 chunk=256):
    """Dense sum over all Gaussians and pixels (cuda/rasterizer.py:74-90), FP64.

    Differentiable when the inputs are FP64 tensors with requires_grad.
    Returns complex128 [C, H, W].
    """
    means = means if isinstance(means, torch.Tensor) and means.dtype == F64 else _t(means)
    inv_cov = inv_cov if isinstance(inv_cov, torch.Tensor) and inv_cov.dtype == F64 else _t(inv_cov)
    colours = colours if isinstance(colours, torch.Tensor) and colours.dtype == F64 else _t(colours)
    phase = phase if isinstance(phase, torch.Tensor) and phase.dtype == F64 else _t(phase)
    opac = opac if isinstance(opac, torch.Tensor) and opac.dtype == F64 else _t(opac)
    curv = curv if isinstance(curv, torch.Tensor) and curv.dtype == F64 else _t(curv)
    area = area if isinstance(area, torch.Tensor) and area.dtype == F64 else _t(area)
    k = _t(k)
    C = colours.shape[1]
    ys, xs = torch.meshgrid(torch.arange(height, dtype=F64), torch.arange(width, dtype=F64),
                            indexing="ij")
    px = xs.reshape(1, -1)
    py = ys.reshape(1, -1)
    out = torch.zeros(C, height * width, dtype=torch.complex128)
    N = means.shape[0]
    for g0 in range(0, N, chunk):
        g1 = min(N, g0 + chunk)
        dx = means[g0:g1, 0:1] - px
        dy = means[g0:g1, 1:2] - py
        mahal, raw, G, alpha, base = pair_terms(
            dx, dy, inv_cov[g0:g1, 0:1], inv_cov[g0:g1, 1:2], inv_cov[g0:g1, 2:3],
            opac[g0:g1, None], curv[g0:g1, None], area[g0:g1, None], dxy2)
        for c in range(C):
            th = phase[g0:g1, c:c + 1] + k[c] * base
            amp = colours[g0:g1, c:c + 1] * alpha
            out[c] = out[c] + torch.complex(amp * torch.cos(th), amp * torch.sin(th)).sum(0)
    return out.view(C, height, width)


def _pairs(bins, width, height):
    """Expand the sorted key list into (pixel x, pixel y, gaussian id, tile) pairs."""
    gx, _ = binning.tile_grid(width, height)
    P = bins["P"]
    if P == 0:
        z = torch.zeros(0, dtype=torch.long)
        return z, z, z
    pl = torch.as_tensor(bins["point_list"][:P].astype(np.int64))
    rng = bins["ranges"]
    tile_of = np.repeat(np.arange(rng.shape[0]), (rng[:, 1] - rng[:, 0]).astype(np.int64))
    tile_of = torch.as_tensor(tile_of)
    ty, tx = tile_of // gx, tile_of % gx
    ly, lx = torch.meshgrid(torch.arange(16), torch.arange(16), indexing="ij")
    pxi = (tx[:, None] * 16 + lx.reshape(1, -1)).reshape(-1)
    pyi = (ty[:, None] * 16 + ly.reshape(1, -1)).reshape(-1)
    gi = pl[:, None].expand(-1, 256).reshape(-1)
    inside = (pxi < width) & (pyi < height)
    return pxi[inside], pyi[inside], gi[inside]


def render_binned(means, inv_cov, colours, phase, opac, curv, area, radii, k, dxy2,
                  width, height, bins=None):
    """Binned forward (cuda/forward.cu:62-162) in FP64: only pairs inside the
    5-sigma tile rectangles contribute.  Returns (complex128 [C,H,W], bins)."""
    if bins is None:
        bins = binning.bin_gaussians(np.asarray(_t(means, torch.float32)), np.asarray(radii),
                                     width, height)
    means, inv_cov, colours, phase = _t(means), _t(inv_cov), _t(colours), _t(phase)
    opac, curv, area, k = _t(opac), _t(curv), _t(area), _t(k)
    C = colours.shape[1]
    pxi, pyi, gi = _pairs(bins, width, height)
    out = torch.zeros(C, height * width, dtype=torch.complex128)
    if gi.numel():
        dx = means[gi, 0] - pxi.to(F64)
        dy = means[gi, 1] - pyi.to(F64)
        _, _, _, alpha, base = pair_terms(dx, dy, inv_cov[gi, 0], inv_cov[gi, 1], inv_cov[gi, 2],
                                          opac[gi], curv[gi], area[gi], dxy2)
        pid = pyi * width + pxi
        for c in range(C):
            th = phase[gi, c] + k[c] * base
            amp = colours[gi, c] * alpha
            out[c].index_add_(0, pid, torch.complex(amp * torch.cos(th), amp * torch.sin(th)))
    return out.view(C, height, width), bins


def render_backward(grad, means, inv_cov, colours, phase, opac, curv, area, k, dxy2,
                    width, height, bins):
    """Closed-form adjoint (cuda/backward.cu:101-148) over the binned pairs, FP64.

    ``grad`` is complex [C,H,W] = dL/dRe U + i dL/dIm U.
    Returns (d_means [N,2], d_inv [N,3], d_col [N,C], d_ph [N,C], d_op, d_cv, d_ar).
    """
    means, inv_cov, colours, phase = _t(means), _t(inv_cov), _t(colours), _t(phase)
    opac, curv, area, k = _t(opac), _t(curv), _t(area), _t(k)
    g = _t(grad, torch.complex128) if not (isinstance(grad, torch.Tensor) and grad.is_complex()) \
        else grad.detach().to(torch.complex128)
    N, C = colours.shape
    d_means = torch.zeros(N, 2, dtype=F64)
    d_inv = torch.zeros(N, 3, dtype=F64)
    d_col = torch.zeros(N, C, dtype=F64)
    d_ph = torch.zeros(N, C, dtype=F64)
    d_op = torch.zeros(N, dtype=F64)
    d_cv = torch.zeros(N, dtype=F64)
    d_ar = torch.zeros(N, dtype=F64)
    pxi, pyi, gi = _pairs(bins, width, height)
    if gi.numel() == 0:
        return d_means, d_inv, d_col, d_ph, d_op, d_cv, d_ar
    dx = means[gi, 0] - pxi.to(F64)
    dy = means[gi, 1] - pyi.to(F64)
    i00, i01, i11 = inv_cov[gi, 0], inv_cov[gi, 1], inv_cov[gi, 2]
    o, cv, ar = opac[gi], curv[gi], area[gi]
    mahal, raw, G, alpha, base = pair_terms(dx, dy, i00, i01, i11, o, cv, ar, dxy2)
    pid = pyi * width + pxi
    gflat = g.reshape(C, -1)
    S_alpha = torch.zeros_like(dx)
    S_base = torch.zeros_like(dx)
    for c in range(C):
        th = phase[gi, c] + k[c] * base
        s, co = torch.sin(th), torch.cos(th)
        gr, gim = gflat[c].real[pid], gflat[c].imag[pid]
        proj_c = gr * co + gim * s
        proj_s = -gr * s + gim * co
        A = colours[gi, c]
        dcol = alpha * proj_c
        dph = A * alpha * proj_s
        S_alpha = S_alpha + A * proj_c
        S_base = S_base + k[c] * dph
        d_col[:, c].index_add_(0, gi, dcol)
        d_ph[:, c].index_add_(0, gi, dph)
    dpow = torch.where(raw >= -50.0, torch.full_like(raw, -0.5), torch.zeros_like(raw))
    dL_dm = S_alpha * o * G * dpow + S_base * 0.5 * cv * ar * dxy2
    d_cv.index_add_(0, gi, S_base * 0.5 * mahal * ar * dxy2)
    d_ar.index_add_(0, gi, S_base * 0.5 * cv * mahal * dxy2)
    d_op.index_add_(0, gi, S_alpha * G)
    d_inv[:, 0].index_add_(0, gi, dL_dm * dx * dx)
    d_inv[:, 1].index_add_(0, gi, dL_dm * 2.0 * dx * dy)
    d_inv[:, 2].index_add_(0, gi, dL_dm * dy * dy)
    d_means[:, 0].index_add_(0, gi, dL_dm * 2.0 * (i00 * dx + i01 * dy))
    d_means[:, 1].index_add_(0, gi, dL_dm * 2.0 * (i11 * dy + i01 * dx))
    return d_means, d_inv, d_col, d_ph, d_op, d_cv, d_ar


def render_binned_window(means, inv_cov, colours, phase, opac, curv, area, k, dxy2,
                         width, height, bins, x0, y0, w, h, chunk_tiles=64):
    """The binned forward (cuda/forward.cu:62-162) in FP64 on the window
    [y0, y0 + h) x [x0, x0 + w) of a width x height frame, from the frame's key list
    ``bins`` (the tiles overlapping the window are expanded, a few at a time, so a
    1080p frame with 150k Gaussians fits in memory).  Returns complex128 [C, h, w]."""
    means, inv_cov, colours, phase = _t(means), _t(inv_cov), _t(colours), _t(phase)
    opac, curv, area, k = _t(opac), _t(curv), _t(area), _t(k)
    C = colours.shape[1]
    gx, gy = binning.tile_grid(width, height)
    rng = bins["ranges"]
    pl = bins["point_list"]
    tiles = [ty * gx + tx for ty in range(y0 // 16, min(gy, (y0 + h + 15) // 16))
             for tx in range(x0 // 16, min(gx, (x0 + w + 15) // 16))]
    out = torch.zeros(C, h * w, dtype=torch.complex128)
    ly, lx = torch.meshgrid(torch.arange(16), torch.arange(16), indexing="ij")
    ly, lx = ly.reshape(1, -1), lx.reshape(1, -1)
    for t0 in range(0, len(tiles), chunk_tiles):
        ts = tiles[t0:t0 + chunk_tiles]
        cnt = np.array([rng[t, 1] - rng[t, 0] for t in ts], dtype=np.int64)
        if cnt.sum() == 0:
            continue
        ids = torch.as_tensor(np.concatenate([pl[rng[t, 0]:rng[t, 1]] for t in ts])
                              .astype(np.int64))
        tile_of = torch.as_tensor(np.repeat(np.array(ts, dtype=np.int64), cnt))
        pxi = ((tile_of % gx)[:, None] * 16 + lx).reshape(-1)
        pyi = ((tile_of // gx)[:, None] * 16 + ly).reshape(-1)
        gi = ids[:, None].expand(-1, 256).reshape(-1)
        keep = (pxi >= x0) & (pxi < min(x0 + w, width)) & (pyi >= y0) & (pyi < min(y0 + h, height))
        pxi, pyi, gi = pxi[keep], pyi[keep], gi[keep]
        dx = means[gi, 0] - pxi.to(F64)
        dy = means[gi, 1] - pyi.to(F64)
        _, _, _, alpha, base = pair_terms(dx, dy, inv_cov[gi, 0], inv_cov[gi, 1],
                                          inv_cov[gi, 2], opac[gi], curv[gi], area[gi], dxy2)
        pid = (pyi - y0) * w + (pxi - x0)
        for c in range(C):
            th = phase[gi, c] + k[c] * base
            amp = colours[gi, c] * alpha
            out[c].index_add_(0, pid, torch.complex(amp * torch.cos(th), amp * torch.sin(th)))
    return out.view(C, h, w)
