"""Seeded synthetic workloads (scenes, targets, depth) and the workload table.

Every array is drawn with ``numpy.random.default_rng(seed)`` in FP64 and cast to
FP32, so the GPU path and the CPU oracle see identical inputs.

Scene recipe (activated values):
  means      uniform over the image
  scales     log-uniform in [1, 8] px per axis
  rotation   uniform [0, pi)
  amplitude  uniform [0, 1] per channel            (``colours``)
  phase      uniform [0, 2 pi) per channel
  opacity    sigmoid(-0.5) ~ 0.3775, the reference's initial value
             (model_2d_gaussian_rgb.py:70)
  curvature  kappa ~ N(0, sigma_kappa) rad/px^2, mapped to the reference's curvature
             c = 2 kappa / (k_ref dxy^2) with k_ref = 2 pi / lambda_min; the raw
             parameter is atanh(c / c_max) (clamped to +-0.999) for the scale-aware
             bound c_max of model_2d_gaussian_rgb.py:103-105.
Target: 40 random anisotropic Gaussian blobs with per-channel amplitudes plus 0.2 x
  (uniform noise smoothed with sigma = 4 px, rescaled to [0, 1]); the sum is rescaled
  to [0, 1].
Depth: uniform noise smoothed with sigma = 16 px, rescaled to [0, 1].
"""
import functools
import math
from argparse import Namespace

import numpy as np

WORKLOADS = {
    # name: (W, H, N, wavelengths, pixel pitch, num_planes, d_val, volume_depth, kappa sigma)
    "gray256_10k": dict(W=256, H=256, N=10_000, wavelengths=[515e-9], pitch=8e-6,
                        distances=[0.0, 2.5e-3, 5e-3], kappa=0.02),
    "rgb1080p_150k": dict(W=1920, H=1080, N=150_000, wavelengths=[639e-9, 515e-9, 473e-9],
                          pitch=8e-6, num_planes=4, d_val=3e-3, volume_depth=6e-3, kappa=0.02),
    "rgb4k_600k": dict(W=3840, H=2160, N=600_000, wavelengths=[639e-9, 515e-9, 473e-9],
                       pitch=8e-6, num_planes=8, d_val=5e-3, volume_depth=10e-3, kappa=0.05),
}


BATCH_RESOLUTIONS = [(512, 512), (1024, 768), (1280, 720), (1920, 1080)]


def batch_specs(n_scenes=16, seed=0, resolutions=None, density=0.07, n_range=(20_000, 150_000),
                num_planes=3, z_range=(0.0, 6e-3)):
    """Specs of a batch of independent RGB scenes of uneven size: resolution drawn from
    ``resolutions``, N = density x pixels clamped to ``n_range``, ``num_planes`` plane
    distances drawn uniformly in ``z_range`` (sorted) per scene.  Scene i uses seed
    ``seed + i`` for its Gaussians, target and depth."""
    rng = np.random.default_rng(seed + 7_000_001)
    res = list(resolutions or BATCH_RESOLUTIONS)
    out = []
    for i in range(n_scenes):
        W, H = res[int(rng.integers(len(res)))]
        N = int(min(max(round(density * W * H), n_range[0]), n_range[1]))
        z = np.sort(rng.uniform(z_range[0], z_range[1], num_planes))
        out.append(dict(W=W, H=H, N=N, wavelengths=[639e-9, 515e-9, 473e-9], pitch=8e-6,
                        distances=[float(np.float32(v)) for v in z], kappa=0.02,
                        seed=seed + i))
    return out


def pad_size_for(W, H):
    """Guard band >= 144 px per side rounded up to 64 (train_2d_gaussian_rgb.py:340-343)."""
    return [((H + 2 * 144 + 63) // 64) * 64, ((W + 2 * 144 + 63) // 64) * 64]


def distances_for(spec):
    """Plane distances as the reference computes them: float32 linspace centred on
    d_val (train_2d_gaussian_rgb.py:359-363), or the explicit list."""
    if "distances" in spec:
        return [float(z) for z in spec["distances"]]
    n = spec["num_planes"]
    if n == 1:
        return [float(spec["d_val"])]
    v = np.linspace(-spec["volume_depth"] / 2.0, spec["volume_depth"] / 2.0, n, dtype=np.float32)
    return [float(np.float32(x) + np.float32(spec["d_val"])) for x in v]


def make_args_prop(spec, renderer="hip"):
    """Namespace with the fields the reference's training script builds (train:335-356)."""
    W, H = spec["W"], spec["H"]
    d = distances_for(spec)
    return Namespace(wavelengths=list(spec["wavelengths"]), pixel_pitch=spec["pitch"],
                     pad_size=pad_size_for(W, H), aperture_size=10 * max(W, H),
                     num_planes=len(d), distances=d, split_ratio=1.0, curv_nyquist=0.9,
                     tile_size=64, gauss_batch=1024, phase_iso=False, grad_ckpt=False,
                     renderer=renderer)


def make_scene(n, W, H, wavelengths, pitch, kappa_sigma=0.02, seed=0, curv_nyquist=0.9):
    """Pre-activation parameters (the seven tensors of Gaussians2D) as FP32 numpy."""
    rng = np.random.default_rng(seed)
    C = len(wavelengths)
    mx = rng.uniform(0, W, n)
    my = rng.uniform(0, H, n)
    s = np.exp(rng.uniform(0.0, math.log(8.0), (n, 2)))
    rot = rng.uniform(0, math.pi, n)
    amp = rng.uniform(0, 1, (n, C))
    ph = rng.uniform(0, 2 * math.pi, (n, C))
    kap = rng.normal(0, kappa_sigma, n)
    lam = float(wavelengths[0]) if C == 1 else float(min(wavelengths))
    k_ref = 2 * math.pi / lam
    c = 2 * kap / (k_ref * pitch ** 2)
    sig = (s + 0.0).max(1)
    cmax = curv_nyquist * lam / (6.0 * sig * pitch ** 2 + 1e-20)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return dict(
        means_2d=f32(np.arctanh(np.clip(np.stack([2 * mx / W - 1, 2 * my / H - 1], 1), -0.999, 0.999))),
        pre_act_scales=f32(np.log(s - 0.1)),
        pre_act_rotation=f32(rot),
        colours=f32(amp),
        pre_act_phase=f32(ph),
        pre_act_opacities=f32(np.full(n, -0.5)),
        curvature=f32(np.arctanh(np.clip(c / cmax, -0.999, 0.999))),
    )


def _smooth_noise(rng, shape, sigma):
    from scipy.ndimage import gaussian_filter
    v = gaussian_filter(rng.uniform(0, 1, shape), sigma, mode="reflect")
    v -= v.min()
    return v / max(v.max(), 1e-12)


def make_target(C, W, H, seed=0, blobs=40):
    """Target intensity [C, H, W] in [0, 1] (FP32)."""
    rng = np.random.default_rng(seed + 1_000_003)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    img = np.zeros((C, H, W))
    for _ in range(blobs):
        cx, cy = rng.uniform(0, W), rng.uniform(0, H)
        sx, sy = rng.uniform(0.02, 0.15) * W, rng.uniform(0.02, 0.15) * H
        th = rng.uniform(0, math.pi)
        a = rng.uniform(0, 1, C)
        dx, dy = xx - cx, yy - cy
        u = (dx * math.cos(th) + dy * math.sin(th)) / sx
        v = (-dx * math.sin(th) + dy * math.cos(th)) / sy
        g = np.exp(-0.5 * (u * u + v * v))
        img += a[:, None, None] * g[None]
    img = img / max(img.max(), 1e-12) + 0.2 * np.stack([_smooth_noise(rng, (H, W), 4.0) for _ in range(C)])
    img -= img.min()
    img /= max(img.max(), 1e-12)
    return img.astype(np.float32)


@functools.lru_cache(maxsize=2)
def frame_targets(C, W, H, seed=0):
    """(make_target, make_depth) of one frame, cached so that several windows of the same
    frame are cut from one synthesis."""
    return make_target(C, W, H, seed), make_depth(W, H, seed)


def make_depth(W, H, seed=0):
    """Depth [1, H, W] in [0, 1] (FP32)."""
    rng = np.random.default_rng(seed + 2_000_029)
    return _smooth_noise(rng, (H, W), 16.0)[None].astype(np.float32)
