"""GPU parity of the fused pad -> glue -> BLASM -> crop kernels against the FP64 oracle.

Bars (relative L2, FP32 kernels vs FP64 oracle with an FP64 transfer-function phase):
  * per-plane intensities <= 1.5e-5 (oracle.model.soft_window_pad -> oracle.model.glue ->
    oracle.optics.BLASM -> crop, i.e. model_2d_gaussian_rgb.py:186-241,
    train_2d_gaussian_rgb.py:89-95,193, utils/propagator.py:248-290);
  * gradient w.r.t. the raster field <= 1.5e-5 against torch autograd of the same FP64
    pipeline;
  * at full 1080p size, the z = 0 plane equals |glue(field)|^2 (a size-independent
    identity of the transform pair).
FFT sizes cover radices 2/4/8, 3, 5, 7, 11 and 13 on runtime plans, Bluestein (chirp-z)
transforms for lengths with the prime factors 19, 23, 59, 73 and 101 (up to a 2432 x 2432
grid), and the compile-time plans of the three standard workloads (transform lengths
1152; 2816 x 4480; 4992 x 8320) at their real sizes.
"""
import numpy as np
import pytest
import torch

from oracle import model as omodel
from oracle import optics as ooptics

pytestmark = pytest.mark.gpu
DEV = "cuda"


def make_field(C, H, W, seed, peak=1.3):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    f = np.zeros((C, H, W), np.complex128)
    for _ in range(12):
        cx, cy, s = rng.uniform(0, W), rng.uniform(0, H), rng.uniform(2, 8)
        a = rng.uniform(0.2, 1.0, C) * np.exp(1j * rng.uniform(0, 2 * np.pi, C))
        g = np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s))
        f += a[:, None, None] * g[None]
    f *= peak / np.abs(f).max()
    f += 0.05 * (rng.normal(size=f.shape) + 1j * rng.normal(size=f.shape))
    return f.astype(np.complex64)


def oracle_planes(field, pad, wl, pitch, dists, aperture, grid, requires_grad=False):
    h = torch.as_tensor(field).to(torch.complex128)
    if requires_grad:
        h.requires_grad_()
    padded = omodel.soft_window_pad(h, pad)
    u, _, _ = omodel.glue(padded)
    prop = ooptics.BLASM(pad, wl, pitch, dists, aperture, grid)
    inten = prop.reconstruct(u)
    H, W = field.shape[1:]
    return ooptics.crop_center(inten, (H, W)), h


def rel(a, b):
    a = torch.as_tensor(a).cpu().to(torch.float64 if not torch.as_tensor(a).is_complex() else torch.complex128)
    b = torch.as_tensor(b).cpu().to(a.dtype)
    return float((a - b).abs().norm() / b.abs().norm())


CASES = [  # (C, H, W, pH, pW, distances, grid)
    (3, 40, 48, 112, 130, [0.0, 2.5e-3, 5e-3], "odak"),      # 224 = 2^5 7, 260 = 2^2 5 13
    (1, 64, 64, 352, 320, [1e-3, 4e-3], "odak"),             # 704 = 2^6 11, 640 = 2^7 5
    (3, 32, 32, 320, 320, [0.0, 6e-3], "fft"),               # the default guard band for 32x32
    (1, 56, 40, 144, 96, [3e-3], "odak"),                     # 288 = 2^5 9, 192 = 2^6 3
]


# Transform lengths with a prime factor above 17 run as Bluestein (chirp-z) transforms
BLUESTEIN_CASES = [
    (3, 40, 50, 46, 118, [0.0, 3e-3], "odak"),            # 92 = 2^2 23, 236 = 2^2 59
    (1, 100, 60, 202, 146, [1e-3], "fft"),                # 404 = 2^2 101, 292 = 2^2 73
    (3, 64, 928, 384, 1216, [2e-3, 5e-3], "odak"),        # rows 2432 = 2^7 19 only
    (1, 928, 928, 1216, 1216, [0.0, 4e-3], "odak"),       # 2432 x 2432
]


@pytest.mark.parametrize("C,H,W,pH,pW,dists,grid", CASES + BLUESTEIN_CASES,
                         ids=[f"{c[3]}x{c[4]}" for c in CASES + BLUESTEIN_CASES])
def test_forward_and_adjoint_match_fp64(pkg, C, H, W, pH, pW, dists, grid):
    from cvqpg_amd.optics import HologramOptics, propagate
    wl = [639e-9, 515e-9, 473e-9][:C] if C == 3 else [515e-9]
    pitch = 8e-6
    ap = 10 * max(W, H) * 10  # open aperture, as the training script configures it
    field = make_field(C, H, W, seed=pH + pW)
    ref, h = oracle_planes(field, (pH, pW), wl, pitch, dists, ap, grid, requires_grad=True)
    opt = HologramOptics((W, H), (pH, pW), wl, pitch, dists, aperture_size=ap, grid=grid)
    f_dev = torch.as_tensor(field).to(DEV).requires_grad_()
    inten = propagate(f_dev, opt)
    torch.cuda.synchronize()
    errs = [rel(inten[d].detach(), ref[d].detach()) for d in range(len(dists))]
    print(f"case C={C} {H}x{W} pad {pH}x{pW} grid={grid}: plane rel L2 "
          + " ".join(f"{e:.1e}" for e in errs), flush=True)
    assert max(errs) <= 1.5e-5
    w = torch.as_tensor(np.random.default_rng(1).uniform(0.5, 1.5, ref.shape))
    (ref * w).sum().backward()
    (inten * w.to(DEV, torch.float32)).sum().backward()
    eg = rel(f_dev.grad, h.grad)
    print(f"  gradient rel L2 {eg:.1e}", flush=True)
    assert eg <= 1.5e-5


FIXED_CASES = [  # (C, H, W, pH, pW, distances): the guard bands of the standard workloads
    (1, 256, 256, 576, 576, [0.0, 2.5e-3, 5e-3]),      # 1152 rows and columns
    (3, 1080, 1920, 1408, 2240, [0.0, 6e-3]),           # 2816 columns, 4480 rows
    (1, 2160, 3840, 2496, 4160, [5e-3]),                # 4992 columns, 8320 rows
    (3, 512, 512, 832, 832, [3e-3]),                    # 1664 rows and columns
    (3, 768, 1024, 1088, 1344, [3e-3]),                 # 2176 columns, 2688 rows
    (3, 720, 1280, 1024, 1600, [3e-3]),                 # 2048 columns, 3200 rows
]


@pytest.mark.parametrize("C,H,W,pH,pW,dists", FIXED_CASES,
                         ids=["256sq", "1080p", "2160p", "512sq", "1024x768", "720p"])
def test_fixed_plan_sizes_match_fp64(pkg, C, H, W, pH, pW, dists):
    """Full-size forward and adjoint on the compile-time transform plans."""
    test_forward_and_adjoint_match_fp64(pkg, C, H, W, pH, pW, dists, "odak")


def test_aperture_and_band_limit_active(pkg):
    """A finite Fourier aperture (export-time setting) and a grid small enough for the
    band limit to bite."""
    from cvqpg_amd.optics import HologramOptics
    C, H, W, pH, pW = 1, 32, 32, 64, 64
    wl, pitch, dists = [515e-9], 8e-6, [2e-2]
    for ap in (40.0, 1e6):
        field = make_field(C, H, W, seed=3, peak=0.9)
        ref, _ = oracle_planes(field, (pH, pW), wl, pitch, dists, ap, "odak")
        opt = HologramOptics((W, H), (pH, pW), wl, pitch, dists, aperture_size=ap)
        inten, _ = opt.forward(torch.as_tensor(field).to(DEV))
        e = rel(inten, ref)
        print(f"aperture {ap}: rel L2 {e:.1e}", flush=True)
        assert e <= 1e-5


def test_reconstruct_full_plane(pkg):
    """propagator.reconstruct interface: phase/amplitude on the padded plane."""
    from cvqpg_amd.optics import HologramOptics
    C, pH, pW = 3, 96, 128
    wl, pitch, dists = [639e-9, 515e-9, 473e-9], 8e-6, [1e-3, 3e-3]
    rng = np.random.default_rng(7)
    ph = rng.uniform(0, 2 * np.pi, (C, pH, pW)).astype(np.float32)
    am = rng.uniform(0, 1, (C, pH, pW)).astype(np.float32)
    prop = ooptics.BLASM((pH, pW), wl, pitch, dists, 1e6, "odak")
    u = torch.polar(torch.as_tensor(am, dtype=torch.float64), torch.as_tensor(ph, dtype=torch.float64))
    ref = prop.reconstruct(u)
    opt = HologramOptics((pW, pH), (pH, pW), wl, pitch, dists, aperture_size=1e6)
    got = opt.reconstruct_planes(torch.as_tensor(ph).to(DEV), torch.as_tensor(am).to(DEV))
    e = rel(got, ref)
    print(f"reconstruct rel L2 {e:.1e}", flush=True)
    assert e <= 1e-5


def test_bluestein_z0_identity_and_linearity(pkg):
    """2432 x 2432 (Bluestein on both axes): the z = 0 plane is |glue(field)|^2, and the
    planes' complex fields are linear in the glued input (field a + 2 i field b)."""
    from cvqpg_amd.optics import HologramOptics
    C, H, W, pH, pW = 1, 928, 928, 1216, 1216
    rng = np.random.default_rng(4)
    opt = HologramOptics((W, H), (pH, pW), [515e-9], 8e-6, [0.0, 3e-3], aperture_size=10 * W)
    fa = (0.3 * (rng.normal(size=(C, H, W)) + 1j * rng.normal(size=(C, H, W)))).astype(np.complex64)
    fb = (0.3 * (rng.normal(size=(C, H, W)) + 1j * rng.normal(size=(C, H, W)))).astype(np.complex64)
    fa, fb = fa / (3 * np.abs(fa).max()), fb / (3 * np.abs(fb).max())   # |.| < 1: glue is identity
    outs = []
    for f in (fa, fb, fa + 2j * fb):
        inten, field = opt.forward(torch.as_tensor(f).to(DEV), keep_field=True)
        outs.append((inten.clone(), field.clone()))
    torch.cuda.synchronize()
    a = torch.as_tensor(np.abs(fa) ** 2).to(DEV)
    e0 = float((outs[0][0][0] - a).norm() / a.norm())
    lin = outs[0][1] + 2j * outs[1][1]
    el = float((outs[2][1] - lin).abs().norm() / lin.abs().norm())
    print(f"2432^2 Bluestein: z=0 identity rel L2 {e0:.1e}, linearity rel L2 {el:.1e}", flush=True)
    assert e0 <= 1e-5 and el <= 1e-5


def test_full_size_z0_identity(pkg):
    """1080p RGB with the training guard band: the z = 0 plane is |glue(field)|^2."""
    from cvqpg_amd.optics import HologramOptics
    C, H, W = 3, 1080, 1920
    pH, pW = 1408, 2240
    rng = np.random.default_rng(0)
    f = (rng.normal(size=(C, H, W)) + 1j * rng.normal(size=(C, H, W))) * 0.5
    field = torch.as_tensor(f.astype(np.complex64)).to(DEV)
    opt = HologramOptics((W, H), (pH, pW), [639e-9, 515e-9, 473e-9], 8e-6, [0.0, 2e-3],
                         aperture_size=10 * W)
    inten, _ = opt.forward(field, keep_field=False)
    a = field.abs()
    expect = torch.where(a > 1, torch.ones_like(a), a) ** 2
    e = float((inten[0] - expect).norm() / expect.norm())
    print(f"full-size z=0 identity rel L2 {e:.1e}", flush=True)
    assert e <= 1e-5
    assert torch.isfinite(inten).all()


@pytest.mark.parametrize("C,H,W,pH,pW,dists", [FIXED_CASES[0], FIXED_CASES[1], CASES[0][:6],
                                                BLUESTEIN_CASES[0][:6]],
                         ids=["1152", "2816x4480", "224x260", "bluestein-92x236"])
def test_tables_built_once_then_kept(pkg, C, H, W, pH, pW, dists):
    """CVQPG_BLASM_TABLES alone builds every transform table; passes with
    CVQPG_BLASM_TABLES_KEPT then reuse them (no table launches) and give bit-identical
    planes, plane fields and field gradients to the plain calls, on a workspace that held
    NaN before the tables-only call (compile-time plans at 1152 and 2816 x 4480, a runtime
    plan and a Bluestein length)."""
    from cvqpg_amd import _lib
    from cvqpg_amd.optics import HologramOptics
    L = _lib.lib()
    wl = [639e-9, 515e-9, 473e-9][:C] if C == 3 else [515e-9]
    opt = HologramOptics((W, H), (pH, pW), wl, 8e-6, dists, aperture_size=10 * max(W, H))
    field = torch.as_tensor(make_field(C, H, W, seed=5)).to(DEV)
    g = torch.rand(len(dists), C, H, W, device=DEV)
    inten, out = opt.forward(field, keep_field=True)
    grad = opt.backward(field, g, out)
    ws = opt._workspace(field.device)
    ws.view(torch.float32)[:ws.numel() // 4].fill_(float("nan"))
    s = _lib.stream_handle(field.device)
    fr = lambda t: _lib.ptr(torch.view_as_real(t))
    _lib.check(L.cvqpg_blasm_forward_passes(8, None, *opt._common(), _lib.ptr(ws), ws.numel(),
                                            None, None, s))
    inten2 = torch.empty_like(inten)
    out2 = torch.empty_like(out)
    grad2 = torch.empty_like(grad)
    for m in (1, 2, 4):
        _lib.check(L.cvqpg_blasm_forward_passes(m | 16, fr(field), *opt._common(), _lib.ptr(ws),
                                                ws.numel(), fr(out2), _lib.ptr(inten2), s))
    _lib.check(L.cvqpg_blasm_backward_passes(7 | 16, fr(field), _lib.ptr(g), fr(out2),
                                             *opt._common(), _lib.ptr(ws), ws.numel(),
                                             fr(grad2), s))
    torch.cuda.synchronize()
    assert torch.equal(inten, inten2) and torch.equal(out, out2) and torch.equal(grad, grad2)


def test_row_pair_layout_with_runtime_row_plan(pkg):
    """A 2816-row grid whose 320-point rows run on the runtime-plan row kernels: the adjoint
    row spectra are stored as row pairs for the prefetching column kernel, written here by
    the runtime adjoint row pass.  Forward and adjoint vs FP64, and the full-plane
    reconstruct backward (oh = 1408, the whole centre half) vs FP64 autograd."""
    test_forward_and_adjoint_match_fp64(pkg, 1, 1080, 96, 1408, 160, [0.0, 3e-3], "odak")
    from cvqpg_amd.optics import HologramOptics, reconstruct_planes
    C, pH, pW, zs = 1, 1408, 160, [2e-3]
    rng = np.random.default_rng(9)
    ph = torch.as_tensor(rng.uniform(0, 2 * np.pi, (C, pH, pW)), dtype=torch.float32)
    am = torch.as_tensor(rng.uniform(0.2, 1.0, (C, pH, pW)), dtype=torch.float32)
    wgt = torch.as_tensor(rng.normal(size=(len(zs), C, pH, pW)), dtype=torch.float32)
    opt = HologramOptics((pW, pH), (pH, pW), [515e-9], 8e-6, zs, aperture_size=1e6)
    php, amp = ph.to(DEV).requires_grad_(), am.to(DEV).requires_grad_()
    (reconstruct_planes(php, amp, opt) * wgt.to(DEV)).sum().backward()
    prop = ooptics.BLASM((pH, pW), [515e-9], 8e-6, zs, 1e6, "odak")
    ph64, am64 = ph.double().requires_grad_(), am.double().requires_grad_()
    (prop.reconstruct(torch.polar(am64, ph64)) * wgt.double()).sum().backward()
    eg_p, eg_a = rel(php.grad, ph64.grad), rel(amp.grad, am64.grad)
    print(f"2816 x 320 full-plane reconstruct gradients vs FP64: phase {eg_p:.1e}, "
          f"amplitude {eg_a:.1e}", flush=True)
    assert eg_p <= 2e-5 and eg_a <= 2e-5
