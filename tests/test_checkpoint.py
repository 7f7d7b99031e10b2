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
