"""The .pth checkpoint exchange with the reference (model_2d_gaussian_rgb.py:155-180).

The reference's save_gaussians writes a dict with exactly the keys below (its parameters
moved to the CPU, img_size as given); load_gaussians reads the seven tensors back.  A
checkpoint written by this package must load in the reference and one written the
reference's way (nn.Parameter values) must load here; torch.load runs with
weights_only=True."""
import torch

REFERENCE_KEYS = {"means_2d", "pre_act_scales", "pre_act_rotation", "colours", "pre_act_phase",
                  "pre_act_opacities", "img_size", "curvature"}


def _model(pkg, seed, C=3, n=500, size=(96, 64)):
    from cvqpg_amd import synthetic
    from cvqpg_amd.model import Gaussians2D
    torch.manual_seed(seed)
    spec = dict(W=size[0], H=size[1], N=n, wavelengths=[639e-9, 515e-9, 473e-9][:C] if C == 3
                else [515e-9], pitch=8e-6, distances=[0.0, 2e-3], kappa=0.02)
    return Gaussians2D(n, size, torch.device("cpu"), synthetic.make_args_prop(spec), channels=C)


def test_save_load_round_trip_and_key_set(pkg, tmp_path):
    from cvqpg_amd.model import PARAM_NAMES
    for C in (3, 1):
        a = _model(pkg, 1, C)
        p = tmp_path / f"g{C}.pth"
        a.save_gaussians(str(p))
        sd = torch.load(str(p), map_location="cpu", weights_only=True)
        assert set(sd) == REFERENCE_KEYS
        assert tuple(sd["img_size"]) == (96, 64)
        assert sd["colours"].shape == (500, C) and sd["pre_act_phase"].shape == (500, C)
        b = _model(pkg, 2, C)
        assert not torch.equal(a.means_2d, b.means_2d)
        b.load_gaussians(str(p))
        for n in PARAM_NAMES:
            assert torch.equal(getattr(a, n), getattr(b, n)), n


def test_loads_a_checkpoint_written_the_reference_way(pkg, tmp_path):
    """The reference saves nn.Parameter values (model_2d_gaussian_rgb.py:156-165)."""
    from cvqpg_amd.model import PARAM_NAMES
    a = _model(pkg, 3)
    ref_style = {n: torch.nn.Parameter(getattr(a, n).detach().clone()).cpu() for n in PARAM_NAMES}
    ref_style["img_size"] = (96, 64)
    p = tmp_path / "ref.pth"
    torch.save(ref_style, str(p))
    b = _model(pkg, 4)
    b.load_gaussians(str(p))
    for n in PARAM_NAMES:
        assert torch.equal(getattr(a, n), getattr(b, n)), n


def test_torch_renderer_gradient_checkpointing(pkg):
    """--renderer torch with args_prop.grad_ckpt (model_2d_gaussian_rgb.py:407-416): the
    field and every parameter gradient equal those of the plain tile loop, which is the
    oracle's 3-sigma renderer on the same activations."""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from oracle import model as omodel
    from cvqpg_amd.model import PARAM_NAMES, Scene2D, make_trainable_2d
    out, grads = [], []
    for ckpt in (False, True):
        g = _model(pkg, 5, n=400, size=(150, 90))
        g.args_prop.renderer = "torch"
        g.args_prop.grad_ckpt = ckpt
        make_trainable_2d(g)
        holo = Scene2D(g, g.args_prop).render_unpadded()
        (holo.abs() ** 2).sum().backward()
        out.append(holo.detach())
        grads.append([getattr(g, n).grad.clone() for n in PARAM_NAMES])
        if not ckpt:
            raw = {n: getattr(g, n).detach() for n in PARAM_NAMES}
            ri = omodel.raster_inputs(raw, (150, 90), g.args_prop.wavelengths,
                                      g.args_prop.pixel_pitch)
            ref = omodel.render_tile_torch(ri, 150, 90)
            assert float((holo.detach() - ref).abs().max()) <= 1e-6 * float(ref.abs().max())
    assert torch.equal(out[0], out[1])
    for a, b in zip(*grads):
        assert torch.allclose(a, b, rtol=1e-6, atol=1e-7 * float(a.abs().max()))
