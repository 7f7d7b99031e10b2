"""Pins the CPU restatement of the warm-start network (oracle/nn.py)."""
import numpy as np
import torch

from oracle import nn as onn


def test_conv_matches_bruteforce():
    g = torch.Generator().manual_seed(0)
    for k in (1, 3):
        x = torch.randn(2, 5, 7, 16, generator=g, dtype=torch.float64)
        w = torch.randn(32, k * k * 16, generator=g, dtype=torch.float64)
        b = torch.randn(32, generator=g, dtype=torch.float64)
        y = onn.conv(x, w, b, k).numpy()
        ref = onn.conv_bruteforce(x.numpy(), w.numpy(), b.numpy(), k)
        assert np.abs(y - ref).max() < 1e-10


def test_conv_tap_offsets():
    """A single nonzero tap moves the image by its offset (ky - 1, kx - 1)."""
    x = torch.zeros(1, 4, 4, 16, dtype=torch.float64)
    x[0, 1, 2, 3] = 1.0
    for tap in range(9):
        w = torch.zeros(16, 9 * 16, dtype=torch.float64)
        w[5, tap * 16 + 3] = 1.0
        y = onn.conv(x, w, None, 3)
        dy, dx = tap // 3 - 1, tap % 3 - 1
        nz = torch.nonzero(y[0, :, :, 5]).tolist()
        assert nz == [[1 - dy, 2 - dx]], (tap, nz)


def test_pool_and_upsample_loops():
    g = torch.Generator().manual_seed(1)
    x = torch.randn(2, 5, 7, 3, generator=g, dtype=torch.float64)
    p = onn.pool2(x)
    assert p.shape == (2, 3, 4, 3)
    for yo in range(3):
        for xo in range(4):
            cells = [(y, xx) for y in (2 * yo, 2 * yo + 1) for xx in (2 * xo, 2 * xo + 1)
                     if y < 5 and xx < 7]
            ref = sum(x[:, y, xx] for y, xx in cells) / len(cells)
            assert torch.allclose(p[:, yo, xo], ref)
    u = onn.up2(p, 5, 7)
    for y in range(5):
        for xx in range(7):
            assert torch.equal(u[:, y, xx], p[:, y // 2, xx // 2])


def test_tiles_from_image_layout():
    img = torch.arange(2 * 4 * 20 * 35, dtype=torch.float64).view(2, 4, 20, 35)
    t = onn.tiles_from_image(img)
    assert t.shape == (2, 2, 3, 1024)
    assert t[1, 1, 2, (3 * 16 + 2) * 16 + 1] == img[1, 3, 18, 33]
    assert t[0, 1, 2, (0 * 16 + 5) * 16 + 4] == 0  # row 21 is outside the image


def test_bf16_rounding_is_nearest_even():
    v = torch.tensor([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9], dtype=torch.float64)
    assert onn.bf16(v).tolist() == [1.0, 1.0 + 2 ** -6, 1.0]


def test_network_gradient_matches_central_differences():
    """Autograd of the composed encoder-decoder (unrounded, FP64) against central
    differences on seeded filter and bias entries of every layer."""
    torch.manual_seed(0)
    img = torch.rand(1, 4, 40, 56, dtype=torch.float64)
    params = {n: (w.double(), b.double() + 0.01) for n, (w, b) in onn.init_params(seed=3).items()}
    target = torch.randn(1, 3, 4, onn.OUT_DIM, dtype=torch.float64)
    weight = torch.rand(1, 3, 4, onn.OUT_DIM, dtype=torch.float64)
    L, _, grads = onn.loss_and_grads(img, params, target, weight, 0.5, rounding=False)
    rng = np.random.default_rng(5)
    h = 1e-6
    for name in ("stem", "e0", "e2b", "e4", "d3a", "d1b", "head"):
        w, b = params[name]
        for which, t, gt in (("w", w, grads[name][0]), ("b", b, grads[name][1])):
            idx = tuple(int(rng.integers(s)) for s in t.shape)
            old = float(t[idx])
            t[idx] = old + h
            lp = float((0.5 * weight * (onn.forward(img, params, False) - target) ** 2).sum())
            t[idx] = old - h
            lm = float((0.5 * weight * (onn.forward(img, params, False) - target) ** 2).sum())
            t[idx] = old
            fd = (lp - lm) / (2 * h)
            an = float(gt[idx])
            assert abs(fd - an) <= 1e-6 + 1e-5 * abs(an), (name, which, idx, fd, an)
