"""Pins oracle.loss (utils/propagator.py:361-449, utils/data_utils.py:22-82).

  * the three-term masked L2 against closed forms (constant offset, disjoint masks,
    injected noise);
  * SSIM(x, x) = 1; SSIM of two constant images against (2ab + C1) / (a^2 + b^2 + C1);
  * SSIM against a brute-force evaluation of Wang et al. (2004), eq. 13, with an
    explicit 11 x 11 Gaussian window (sigma 1.5) and pure-Python sums over every valid
    window position, on a tiny image;
  * the gradient of the step loss against FP64 central finite differences;
  * target synthesis: the plane masks partition the image, a single plane returns the
    image itself (sigma 0 is a delta kernel), the blur kernels are normalised (a constant
    image stays constant away from the border), and every plane's target is the blurred
    image restricted to that plane's mask, as a per-pixel loop over the masks.
The defocus kernel's grid (odak's ``generate_2d_gaussian`` on linspace(-L/2, L/2, L)) is
recalled from odak's public source and stays "parity unpinned".
"""
import math

import numpy as np
import torch

from oracle import loss

F64 = torch.float64


def test_plane_loss_closed_forms():
    rng = np.random.default_rng(0)
    T = torch.as_tensor(rng.uniform(0, 1, (2, 9, 11)))
    a = 0.125
    M = torch.ones_like(T)
    got = float(loss.plane_loss(T + a, T, M))
    expect = a * a + a * a + a * a * float((T * T).mean())
    assert abs(got - expect) <= 1e-14
    # a mask of zeros removes the masked term
    got0 = float(loss.plane_loss(T + a, T, torch.zeros_like(T)))
    assert abs(got0 - (a * a + a * a * float((T * T).mean()))) <= 1e-14
    # noise n added to the target in the masked and correlation terms only
    n = torch.as_tensor(rng.normal(0, 1e-3, T.shape))
    P = torch.as_tensor(rng.uniform(0, 1, T.shape))
    got_n = float(loss.plane_loss(P, T, M, n))
    e = (float(((P - T) ** 2).mean()) + float(((P - T - n) ** 2).mean())
         + float(((P * (T + n) - (T + n) ** 2) ** 2).mean()))
    assert abs(got_n - e) <= 1e-14


def test_ssim_identity_and_constants():
    rng = np.random.default_rng(1)
    X = torch.as_tensor(rng.uniform(0, 1, (3, 20, 17)))
    assert abs(float(loss.ssim(X, X)) - 1.0) <= 1e-14
    C1 = (0.01) ** 2
    for a, b in ((0.2, 0.7), (0.0, 1.0), (0.5, 0.5)):
        A = torch.full((1, 13, 12), a, dtype=F64)
        B = torch.full((1, 13, 12), b, dtype=F64)
        expect = (2 * a * b + C1) / (a * a + b * b + C1)
        assert abs(float(loss.ssim(A, B)) - expect) <= 1e-12


def _ssim_brute(X, Y):
    """Wang et al. 2004, eq. 13 with a Gaussian-weighted 11x11 window, pure Python."""
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    g = [math.exp(-((i - 5) ** 2) / (2 * 1.5 ** 2)) for i in range(11)]
    s = sum(g)
    w = [[g[i] * g[j] / (s * s) for j in range(11)] for i in range(11)]
    C, H, W = len(X), len(X[0]), len(X[0][0])
    per_c = []
    for c in range(C):
        vals = []
        for y in range(H - 10):
            for x in range(W - 10):
                mx = my = sxx = syy = sxy = 0.0
                for i in range(11):
                    for j in range(11):
                        a, b, ww = X[c][y + i][x + j], Y[c][y + i][x + j], w[i][j]
                        mx += ww * a
                        my += ww * b
                        sxx += ww * a * a
                        syy += ww * b * b
                        sxy += ww * a * b
                vx, vy, cxy = sxx - mx * mx, syy - my * my, sxy - mx * my
                vals.append((2 * mx * my + C1) * (2 * cxy + C2)
                            / ((mx * mx + my * my + C1) * (vx + vy + C2)))
        per_c.append(sum(vals) / len(vals))
    return sum(per_c) / C


def test_ssim_matches_brute_force_window():
    rng = np.random.default_rng(2)
    X = rng.uniform(0, 1, (2, 14, 13))
    Y = np.clip(X + rng.normal(0, 0.2, X.shape), 0, 1)
    got = float(loss.ssim(torch.as_tensor(X), torch.as_tensor(Y)))
    expect = _ssim_brute(X.tolist(), Y.tolist())
    assert abs(got - expect) <= 1e-12, (got, expect)
    assert 0.0 < expect < 0.99


def test_step_loss_gradient_matches_central_fd():
    rng = np.random.default_rng(3)
    D, C, pH, pW, H, W = 2, 1, 24, 26, 16, 18
    I = torch.as_tensor(rng.uniform(0.05, 0.95, (D, C, pH, pW))).requires_grad_()
    T = torch.as_tensor(rng.uniform(0, 1, (D, C, H, W)))
    M = torch.as_tensor((rng.uniform(0, 1, (D, C, H, W)) > 0.5).astype(np.float64))
    L = loss.step_loss(I, T, M, H, W)
    L.backward()
    g = I.grad.reshape(-1)
    h = 1e-6
    flat = I.detach().reshape(-1)
    # the centre crop rows 4..19, columns 4..21: entries inside and outside it
    for idx in [0, 5 * pW + 4, 10 * pW + 10, pH * pW + 12 * pW + 21, 2 * pH * pW - 1]:
        p, m = flat.clone(), flat.clone()
        p[idx] += h
        m[idx] -= h
        fd = (float(loss.step_loss(p.view_as(I), T, M, H, W))
              - float(loss.step_loss(m.view_as(I), T, M, H, W))) / (2 * h)
        assert abs(float(g[idx]) - fd) <= 1e-7 * max(1.0, abs(fd)) + 1e-10, (idx, float(g[idx]), fd)


def test_targets_partition_and_blur():
    rng = np.random.default_rng(4)
    C, H, W, P = 2, 40, 44, 3
    img = torch.as_tensor(rng.uniform(0, 1, (C, H, W)))
    depth = torch.as_tensor(rng.uniform(0, 1, (1, H, W)))
    targets, masks, q = loss.make_targets(img, depth, P)
    # masks partition the image, by the quantised depth round(depth * (P - 1))
    np.testing.assert_array_equal(masks.sum(0).numpy(), np.ones((C, H, W)))
    dq = np.round(depth.numpy()[0] * (P - 1))
    for d in range(P):
        np.testing.assert_array_equal(masks[d, 0].numpy(), (dq == d).astype(float))
    # one plane: the target is the image (sigma 0 is a delta kernel)
    t1, m1, _ = loss.make_targets(img, depth, 1)
    np.testing.assert_allclose(t1[0].numpy(), img.numpy(), rtol=0, atol=1e-12)
    # a constant image stays constant on every plane away from the border: each plane's
    # target at a pixel is one normalised blur of the whole image
    t_c, m_c, _ = loss.make_targets(torch.full((1, H, W), 0.3, dtype=F64), depth, P)
    inner = t_c[:, 0, 11:-11, 11:-11].numpy()
    np.testing.assert_allclose(inner, 0.3, rtol=0, atol=1e-12)
    # per pixel: target_d = sum_j [mask_j] * blur_{|d-j|}(image)
    for d in range(P):
        acc = np.zeros((C, H, W))
        for j in range(P):
            s = 0.0 if d == j else float(int(abs(d - j) * 8))
            ker = loss.odak_gaussian_2d(21, s)
            ker = ker / ker.sum()
            for c in range(C):
                b = torch.nn.functional.conv2d(img[c][None, None], ker[None, None], padding=10)
                acc[c] += b[0, 0].numpy() * masks[j, c].numpy()
        np.testing.assert_allclose(targets[d].numpy(), acc, rtol=0, atol=1e-12)
