"""Target synthesis of the product loss object against the oracle (CPU).

``multiplane_loss_odak.set_targets`` / ``add_defocus_blur`` (utils/propagator.py:361-415)
and the factory ``multiplane_loss`` (utils/data_utils.py:66-82) build the per-plane targets
and masks every training step is fitted to.  They run once per fit, in PyTorch, so they are
compared here with oracle.loss.make_targets on the same FP32 inputs (the oracle computes in
FP64): masks equal, targets within 1e-6 relative L2.  Cases: the factory values (blur size
20 -> 21 taps, ratio 8) at 1, 2, 3, 4 and 8 planes, the two-plane split ratio, an odd blur
size and a fractional blur ratio whose int() truncation gives sigma 0 (a delta kernel)
between neighbouring planes, and a depth map that leaves a plane without pixels.
"""
from argparse import Namespace

import numpy as np
import pytest
import torch

from oracle import loss as oloss


def inputs(C, H, W, seed, empty_plane=None, P=4):
    rng = np.random.default_rng(seed)
    img = torch.as_tensor(rng.uniform(0, 1, (C, H, W)), dtype=torch.float32)
    depth = torch.as_tensor(rng.uniform(0, 1, (1, H, W)), dtype=torch.float32)
    if empty_plane is not None:   # move every pixel of that plane to a neighbour
        q = torch.round(depth * (P - 1))
        depth = torch.where(q == empty_plane, torch.clamp(depth + 1.0 / (P - 1), max=1.0), depth)
    return img, depth


def rel(a, b):
    return float((a.double() - b.double()).norm() / b.double().norm())


@pytest.mark.parametrize("C,H,W,P,split", [(3, 48, 64, 4, 1.0), (1, 40, 36, 1, 1.0),
                                           (3, 32, 50, 2, 0.6), (1, 33, 47, 3, 1.0),
                                           (3, 40, 40, 8, 1.0)])
def test_factory_targets_match_oracle(pkg, C, H, W, P, split):
    from cvqpg_amd import loss
    img, depth = inputs(C, H, W, P + C)
    args = Namespace(num_planes=P, split_ratio=split)
    targets, fn, masks = loss.multiplane_loss(img, depth, args)
    ref_t, ref_m, q = oloss.make_targets(img.double(), depth.double(), P, blur_size=20,
                                         blur_ratio=8, split_ratio=split)
    assert torch.equal(masks, ref_m.float())
    assert rel(targets, ref_t) <= 1e-6
    _, _, dq = fn.get_targets()
    np.testing.assert_array_equal(dq.numpy(), (q.float() / max(P - 1, 1)).numpy())


@pytest.mark.parametrize("blur_size,blur_ratio", [(11, 0.25), (9, 3), (20, 1.5)])
def test_blur_settings_match_oracle(pkg, blur_size, blur_ratio):
    from cvqpg_amd import loss
    img, depth = inputs(3, 36, 44, 7)
    fn = loss.multiplane_loss_odak(img, depth, blur_ratio=blur_ratio, target_blur_size=blur_size,
                                   number_of_planes=5, backend="torch")
    t, m, _ = fn.get_targets()
    ref_t, ref_m, _ = oloss.make_targets(img.double(), depth.double(), 5, blur_size=blur_size,
                                         blur_ratio=blur_ratio)
    assert torch.equal(m, ref_m.float())
    assert rel(t, ref_t) <= 1e-6


def test_empty_plane(pkg):
    """A depth map that leaves plane 2 without pixels: its mask is empty and it is skipped
    as a blur source (the reference's torch.sum(targets_cache[j]) > 0 test), while its
    target still holds the other planes' pixels blurred by their distance to plane 2."""
    from cvqpg_amd import loss
    P = 4
    img, depth = inputs(1, 40, 40, 3, empty_plane=2, P=P)
    fn = loss.multiplane_loss_odak(img, depth, blur_ratio=8, target_blur_size=20,
                                   number_of_planes=P, backend="torch")
    t, m, _ = fn.get_targets()
    assert float(m[2].sum()) == 0.0 and float(t[2].abs().max()) > 0.0
    ref_t, ref_m, _ = oloss.make_targets(img.double(), depth.double(), P)
    assert torch.equal(m, ref_m.float())
    assert rel(t, ref_t) <= 1e-6
