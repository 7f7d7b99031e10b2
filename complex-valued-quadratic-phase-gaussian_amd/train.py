"""Training-step driver with the reference's structure (train_2d_gaussian_rgb.py).

* ``reconstruct``          -- train_2d_gaussian_rgb.py:89-95
* ``MeansCosineScheduler`` -- :58-70 (closed form of CosineAnnealingLR, eta_min 1e-3)
* ``setup_optimizer``      -- :73-86 (six learning-rate groups, Adan foreach=False)
* ``AutogradTrainer``      -- one iteration of :187-211 through PyTorch autograd:
  render (HIP rasterizer via ``qpf_render``) -> pad -> glue -> propagator ->
  per-plane loss + SSIM -> backward -> nan_to_num + clip -> Adan -> schedule.
  Logging reads (``loss.item()``, the per-step PSNR) are left out of the step so it
  runs without host synchronisation; ``psnr`` computes it on demand.
"""
import copy
import math

import torch

from . import loss as loss_mod
from .model import Gaussians2D, Scene2D, make_trainable_2d, PARAM_NAMES
from .optimizer import Adan
from .propagator import propagator as Propagator
from . import synthetic


def reconstruct(scene, prop, img_size):
    """train_2d_gaussian_rgb.py:89-95."""
    h = scene.render(img_size)
    phase_map = torch.angle(h) % (2 * math.pi)
    amplitude = torch.clamp(torch.abs(h), min=0.0, max=1.0)
    intens = prop.reconstruct(phase_map, amplitude=amplitude, no_grad=False)
    return torch.sum(intens, dim=0), phase_map, amplitude


class MeansCosineScheduler:
    def __init__(self, optimizer, num_itrs, eta_min=1e-3):
        self.optimizer = optimizer
        self.idx = [i for i, g in enumerate(optimizer.param_groups) if g.get("name") == "means"]
        self.base = [optimizer.param_groups[i]["lr"] for i in self.idx]
        self.T = max(1, num_itrs)
        self.eta_min = eta_min
        self.t = 0

    def lr_at(self, t, base):
        return self.eta_min + (base - self.eta_min) * (1 + math.cos(math.pi * t / self.T)) / 2

    def step(self):
        self.t += 1
        for i, b in zip(self.idx, self.base):
            self.optimizer.param_groups[i]["lr"] = self.lr_at(self.t, b)


def setup_optimizer(gaussians, num_itrs, lr=0.01):
    groups = [
        {"params": [gaussians.means_2d], "lr": lr, "name": "means"},
        {"params": [gaussians.pre_act_scales], "lr": 0.005, "name": "scales"},
        {"params": [gaussians.colours, gaussians.pre_act_phase], "lr": 0.0025, "name": "amplitude_phase"},
        {"params": [gaussians.pre_act_opacities], "lr": 0.025, "name": "opacity"},
        {"params": [gaussians.pre_act_rotation], "lr": 0.001, "name": "rotation"},
        {"params": [gaussians.curvature], "lr": 0.001, "name": "curvature"},
    ]
    opt = Adan(groups, lr=0, eps=1e-8, foreach=False)
    return opt, MeansCosineScheduler(opt, num_itrs), [p for g in groups for p in g["params"]]


def psnr(pred, target):
    mse = torch.mean((pred.clamp(0, 1) - target) ** 2)
    return float(10 * torch.log10(1.0 / mse))


class Workload:
    """Scene parameters, targets and optics of one synthetic workload on a device."""

    def __init__(self, name, device="cuda", seed=0, renderer="hip", spec=None):
        spec = dict(synthetic.WORKLOADS[name] if spec is None else spec)
        self.name, self.spec, self.device = name, spec, torch.device(device)
        W, H, N = spec["W"], spec["H"], spec["N"]
        self.img_size = (W, H)
        self.args_prop = synthetic.make_args_prop(spec, renderer)
        C = len(spec["wavelengths"])
        raw = synthetic.make_scene(N, W, H, spec["wavelengths"], spec["pitch"], spec["kappa"], seed)
        self.gaussians = Gaussians2D(N, (W, H), self.device, self.args_prop, channels=C, init=False)
        with torch.no_grad():
            for n in PARAM_NAMES:
                getattr(self.gaussians, n).copy_(torch.as_tensor(raw[n]))
        if "crop" in spec:  # a W x H window of a larger synthetic frame
            x0, y0, Ws, Hs = spec["crop"]
            img, dep = synthetic.frame_targets(C, Ws, Hs, seed)
            img, dep = img[:, y0:y0 + H, x0:x0 + W], dep[:, y0:y0 + H, x0:x0 + W]
        else:
            img, dep = synthetic.make_target(C, W, H, seed), synthetic.make_depth(W, H, seed)
        self.target_image = torch.as_tensor(img.copy()).to(self.device)
        self.target_depth = torch.as_tensor(dep.copy()).to(self.device)
        targets, self.loss_fn, self.masks = loss_mod.multiplane_loss(
            self.target_image, self.target_depth, self.args_prop)
        self.targets = targets

    def make_propagator(self, backend="hip"):
        ap = self.args_prop
        return Propagator(resolution=ap.pad_size, wavelengths=ap.wavelengths,
                          pixel_pitch=ap.pixel_pitch, number_of_frames=len(ap.wavelengths),
                          distances=ap.distances, aperture_size=ap.aperture_size,
                          device=self.device, backend=backend)


class AutogradTrainer:
    """One training iteration through autograd (train_2d_gaussian_rgb.py:187-211)."""

    def __init__(self, wl, num_itrs=2001, backend="torch", curvature_frozen=False):
        """``backend`` selects the propagator and loss implementation: "hip" (the gfx950
        kernels behind the reference's objects) or "torch" (the reference's PyTorch
        structure, the baseline).  The rasterizer follows ``args_prop.renderer``."""
        self.wl = wl
        self.backend = backend
        self.loss_fn = copy.copy(wl.loss_fn)
        self.loss_fn.backend = backend
        self.scene = Scene2D(wl.gaussians, wl.args_prop)
        make_trainable_2d(wl.gaussians)
        if curvature_frozen:
            wl.gaussians.curvature.requires_grad_(False)
        self.prop = wl.make_propagator(backend)
        self.opt, self.sched, self.params = setup_optimizer(wl.gaussians, num_itrs)

    def step(self, noises=None):
        W, H = self.wl.img_size
        self.opt.zero_grad(set_to_none=True)
        intens, _, _ = reconstruct(self.scene, self.prop, self.wl.img_size)
        pH, pW = self.wl.args_prop.pad_size
        total = 0.0
        for d in range(intens.shape[0]):
            pred = intens[d][:, pH // 2 - H // 2:pH // 2 + H // 2, pW // 2 - W // 2:pW // 2 + W // 2]
            pred = torch.clamp(pred, 0.0, 1.0)
            nz = None if noises is None else noises[d]
            total = total + self.loss_fn(pred, self.wl.targets[d], d, inject_noise=True, noise=nz)
            total = total + loss_mod.GaussianLoss(pred, self.wl.targets[d], backend=self.backend)
        total.backward()
        for p in self.params:
            if p.grad is not None:
                p.grad = torch.nan_to_num(p.grad, nan=0.0, posinf=0.0, neginf=0.0)
        torch.nn.utils.clip_grad_norm_(self.params, max_norm=1.0)
        self.opt.step()
        self.sched.step()
        return total.detach()
