"""GPU parity of the reference's own objects on the HIP backend.

``propagator.reconstruct`` (utils/propagator.py:224-290) and the loss objects
``multiplane_loss_odak.__call__`` (utils/propagator.py:417-449) and ``GaussianLoss``
(utils/data_utils.py:22-41) with ``backend='hip'`` run the gfx950 kernels.  Oracles:
oracle.optics.BLASM (FP64 transfer function, torch autograd in FP64 for the gradients),
oracle.loss.plane_loss / ssim_term (FP64).

Bars (measured values in the printed lines): intensities <= 1e-5 relative L2 (FP32 FFTs
of ~10^6 points: ~1e-6), phase / amplitude gradients <= 2e-5 relative L2, loss values
<= 1e-5 relative, loss gradients <= 1e-5 relative L2.
"""
import math

import numpy as np
import pytest
import torch

from oracle import loss as oloss
from oracle import optics as ooptics

pytestmark = pytest.mark.gpu
DEV = "cuda"
WL3 = [639e-9, 515e-9, 473e-9]


def rel(a, b):
    a = a.detach().cpu().double()
    b = b.detach().cpu().double()
    return float((a - b).norm() / b.norm())


def make_prop(pH, pW, zs, backend, wl=WL3, **kw):
    from cvqpg_amd.propagator import propagator
    return propagator(resolution=(pH, pW), wavelengths=wl, pixel_pitch=8e-6,
                      number_of_frames=len(wl), distances=zs, device=DEV, backend=backend,
                      **kw)


def phases_amps(C, pH, pW, seed):
    rng = np.random.default_rng(seed)
    ph = torch.as_tensor(rng.uniform(0, 2 * math.pi, (C, pH, pW)), dtype=torch.float32)
    am = torch.as_tensor(rng.uniform(0.2, 1.0, (C, pH, pW)), dtype=torch.float32)
    return ph, am


@pytest.mark.parametrize("pH,pW,zs", [(128, 192, [0.0, 2e-3]), (384, 448, [1e-3, 4e-3, 6e-3]),
                                      (120, 146, [1e-3, 3e-3])])   # 292 = 2^2 73: Bluestein
def test_reconstruct_hip_matches_fp64_and_torch(pkg, pH, pW, zs):
    ph, am = phases_amps(3, pH, pW, 1)
    hip = make_prop(pH, pW, zs, "hip", aperture_size=10 * max(pH, pW))
    got = hip.reconstruct(ph.to(DEV), am.to(DEV))
    assert got.shape == (3, len(zs), 3, pH, pW)
    torch_p = make_prop(pH, pW, zs, "torch", aperture_size=10 * max(pH, pW))
    tref = torch_p.reconstruct(ph.to(DEV), am.to(DEV))
    bl = ooptics.BLASM((pH, pW), WL3, 8e-6, zs, aperture_size=10 * max(pH, pW))
    u = torch.polar(am.double(), ph.double())
    ref = bl.reconstruct(u)                                     # [D, C, pH, pW]
    for f in range(3):
        for c in range(3):
            if c != f:
                assert float(got[f, :, c].abs().max()) == 0.0
    e = rel(torch.stack([got[c, :, c] for c in range(3)], 1), ref)
    et = rel(torch.stack([tref[c, :, c] for c in range(3)], 1), ref)
    print(f"{pH}x{pW} z={zs}: HIP vs FP64 {e:.1e}, torch backend vs FP64 {et:.1e}", flush=True)
    assert e <= 1e-5


def test_reconstruct_hip_gradients_match_fp64_autograd(pkg):
    pH, pW, zs = 128, 160, [0.0, 3e-3]
    ph, am = phases_amps(3, pH, pW, 2)
    rng = np.random.default_rng(3)
    wgt = torch.as_tensor(rng.normal(size=(len(zs), 3, pH, pW)), dtype=torch.float32)
    hip = make_prop(pH, pW, zs, "hip", aperture_size=10 * max(pH, pW))
    php, amp = ph.to(DEV).requires_grad_(), am.to(DEV).requires_grad_()
    out = hip.reconstruct(php, amp, no_grad=False)
    (torch.stack([out[c, :, c] for c in range(3)], 1) * wgt.to(DEV)).sum().backward()
    bl = ooptics.BLASM((pH, pW), WL3, 8e-6, zs, aperture_size=10 * max(pH, pW))
    ph64, am64 = ph.double().requires_grad_(), am.double().requires_grad_()
    (bl.reconstruct(torch.polar(am64, ph64)) * wgt.double()).sum().backward()
    eg_p, eg_a = rel(php.grad, ph64.grad), rel(amp.grad, am64.grad)
    print(f"reconstruct gradients vs FP64 autograd: phase {eg_p:.1e}, amplitude {eg_a:.1e}",
          flush=True)
    assert eg_p <= 2e-5 and eg_a <= 2e-5


def test_reconstruct_hip_general_channel_power(pkg):
    """One frame lighting all three channels with unequal power (utils/propagator.py:280-286)."""
    pH, pW, zs = 96, 128, [2e-3]
    ph, am = phases_amps(3, pH, pW, 4)
    power = torch.tensor([[1.0, 0.5, 0.0]])
    from cvqpg_amd.propagator import propagator
    kw = dict(resolution=(pH, pW), wavelengths=WL3, pixel_pitch=8e-6, number_of_frames=1,
              distances=zs, device=DEV, laser_channel_power=power, aperture_size=1e5)
    got = propagator(backend="hip", **kw).reconstruct(ph[:1].to(DEV), am.to(DEV))
    ref = propagator(backend="torch", **kw).reconstruct(ph[:1].to(DEV), am.to(DEV))
    assert got.shape == (1, 1, 3, pH, pW)
    assert float(got[0, 0, 2].abs().max()) == 0.0
    e = rel(got, ref)
    print(f"channel power [1, 0.5, 0]: HIP vs torch backend {e:.1e}", flush=True)
    assert e <= 1e-5


def test_custom_aperture_on_both_backends(pkg):
    """A custom aperture tensor at the hologram resolution, zero-padded to the 2x grid and
    applied to the spectrum and to H (utils/propagator.py:59-61,199-200): the torch backend
    and the HIP kernels (cvqpg_blasm_reconstruct_aperture) against the FP64 oracle, on a
    runtime-plan grid and on the 1080p guard band's 2816-row grid (compile-time column
    plans, row-pair adjoint layout), intensities and, through no_grad=False, the phase and
    amplitude gradients."""
    from cvqpg_amd.propagator import propagator
    for pH, pW, zs in ((64, 96, [1e-3, 4e-3]), (1408, 160, [2e-3])):
        ph, am = phases_amps(1, pH, pW, 5)
        rng = np.random.default_rng(6)
        ap = torch.as_tensor(rng.uniform(0.3, 1.0, (pH, pW)), dtype=torch.float32)
        ap[: pH // 4] = 0.0                          # a blocked band
        kw = dict(resolution=(pH, pW), wavelengths=[515e-9], pixel_pitch=8e-6,
                  number_of_frames=1, distances=zs, device=DEV, aperture=ap)
        bl = ooptics.BLASM((pH, pW), [515e-9], 8e-6, zs, aperture=ap.double())
        ph64, am64 = ph.double().requires_grad_(), am.double().requires_grad_()
        ref = bl.reconstruct(torch.polar(am64, ph64))               # [D, 1, pH, pW]
        wgt = torch.as_tensor(rng.normal(size=ref.shape))
        (ref * wgt).sum().backward()
        got_t = propagator(backend="torch", **kw).reconstruct(ph.to(DEV), am.to(DEV))
        php, amp = ph.to(DEV).requires_grad_(), am.to(DEV).requires_grad_()
        got_h = propagator(backend="hip", **kw).reconstruct(php, amp, no_grad=False)
        (got_h[0] * wgt.to(DEV, torch.float32)).sum().backward()
        et, eh = rel(got_t[0], ref), rel(got_h[0], ref)
        eg_p, eg_a = rel(php.grad, ph64.grad), rel(amp.grad, am64.grad)
        print(f"custom aperture {pH}x{pW}: torch backend {et:.1e}, HIP {eh:.1e} vs FP64; "
              f"HIP gradients phase {eg_p:.1e}, amplitude {eg_a:.1e}", flush=True)
        assert et <= 1e-5 and eh <= 1e-5
        assert eg_p <= 2e-5 and eg_a <= 2e-5


@pytest.mark.parametrize("noise", [False, True])
def test_plane_loss_call_matches_oracle(pkg, noise):
    from cvqpg_amd import loss as ploss
    rng = np.random.default_rng(7)
    C, H, W = 3, 70, 90
    img = torch.as_tensor(rng.uniform(0, 1, (C, H, W)), dtype=torch.float32)
    dep = torch.as_tensor(rng.uniform(0, 1, (1, H, W)), dtype=torch.float32)
    fn = ploss.multiplane_loss_odak(img.to(DEV), dep.to(DEV), target_blur_size=20,
                                    number_of_planes=3, blur_ratio=8, weights=[1.0, 1.0, 1.0, 0.0],
                                    device=DEV, backend="hip")
    pred = torch.as_tensor(rng.uniform(0, 1, (C, H, W)), dtype=torch.float32)
    nz = torch.as_tensor(rng.normal(size=(C, H, W)), dtype=torch.float32)
    d = 1
    T = fn.targets[d]
    p = pred.to(DEV).requires_grad_()
    L = fn(p, T, d, inject_noise=noise, noise=nz.to(DEV) if noise else None)
    L.backward()
    T64, M64 = T.cpu().double(), fn.masks[d].cpu().double()
    scaled = nz.double() * 1e-3 * float(T64.max() - T64.min()) if noise else None
    p64 = pred.double().requires_grad_()
    L64 = oloss.plane_loss(p64, T64, M64, scaled)
    L64.backward()
    el = abs(float(L) - float(L64)) / float(L64)
    eg = rel(p.grad, p64.grad)
    print(f"__call__ (noise {noise}): loss {float(L):.6e} vs {float(L64):.6e} ({el:.1e}), "
          f"grad {eg:.1e}", flush=True)
    assert el <= 1e-5 and eg <= 1e-5


def test_plane_loss_internal_noise_draw(pkg):
    from cvqpg_amd import loss as ploss
    rng = np.random.default_rng(8)
    C, H, W = 1, 200, 240
    img = torch.as_tensor(rng.uniform(0, 1, (C, H, W)), dtype=torch.float32).to(DEV)
    dep = torch.as_tensor(rng.uniform(0, 1, (1, H, W)), dtype=torch.float32).to(DEV)
    fn = ploss.multiplane_loss_odak(img, dep, number_of_planes=2, device=DEV, backend="hip")
    z1, z2 = fn._draw(fn.targets[0]), fn._draw(fn.targets[0])
    assert not torch.equal(z1, z2)
    for z in (z1, z2):
        m, s = float(z.mean()), float(z.std())
        assert abs(m) < 0.02 and abs(s - 1) < 0.02, (m, s)
    pred = img.clone()
    a = float(fn(pred, fn.targets[0], 0, inject_noise=False))
    b = float(fn(pred, fn.targets[0], 0, inject_noise=True))
    assert b != a and math.isfinite(b)


def test_gaussian_loss_matches_oracle(pkg):
    from cvqpg_amd import loss as ploss
    rng = np.random.default_rng(9)
    C, H, W = 3, 120, 150
    T = torch.as_tensor(rng.uniform(0, 1, (C, H, W)), dtype=torch.float32)
    P = torch.clamp(T + torch.as_tensor(rng.normal(0, 0.1, (C, H, W)), dtype=torch.float32), 0, 1)
    p = P.to(DEV).requires_grad_()
    L = ploss.GaussianLoss(p, T.to(DEV), backend="hip")
    L.backward()
    p64 = P.double().requires_grad_()
    L64 = oloss.ssim_term(p64, T.double())
    L64.backward()
    el = abs(float(L) - float(L64)) / float(L64)
    eg = rel(p.grad, p64.grad)
    print(f"GaussianLoss: {float(L):.6e} vs {float(L64):.6e} ({el:.1e}), grad {eg:.1e}", flush=True)
    assert el <= 1e-5 and eg <= 1e-5


def test_reference_structured_step_on_hip_matches_torch_backend(pkg):
    """train_2d_gaussian_rgb.py:187-211 with the reference's objects: the HIP backend
    (propagator and loss kernels behind the same objects) against the torch backend, from
    the same start with the same target noise."""
    from cvqpg_amd import train
    spec = dict(W=96, H=80, N=600, wavelengths=WL3, pitch=8e-6, distances=[0.0, 3e-3],
                kappa=0.02)
    res = {}
    for backend in ("torch", "hip"):
        wl = train.Workload("objects", device=DEV, seed=2, spec=spec)
        tr = train.AutogradTrainer(wl, num_itrs=100, backend=backend)
        rng = np.random.default_rng(10)
        noises = [torch.as_tensor(rng.normal(size=(3, 80, 96)), dtype=torch.float32).to(DEV)
                  for _ in range(2)]
        L = tr.step(noises=noises)
        torch.cuda.synchronize()
        res[backend] = (float(L), [p.detach().clone() for p in tr.params])
    (lt, pt), (lh, ph) = res["torch"], res["hip"]
    el = abs(lh - lt) / abs(lt)
    ep = max(rel(a, b) for a, b in zip(ph, pt))
    print(f"reference-structured step: loss hip {lh:.6e} torch {lt:.6e} ({el:.1e}); "
          f"parameters after the step {ep:.1e}", flush=True)
    assert el <= 1e-4 and ep <= 1e-5
