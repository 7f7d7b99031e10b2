"""Pins oracle.model (model_2d_gaussian_rgb.py:83-153, 186-241, 269-439;
train_2d_gaussian_rgb.py:89-95).

  * activations: means, scales, phase, opacity and the scale-aware curvature bound
    against their closed forms (the bound: |c| < 0.9 lambda_min / (6 sigma_max dxy^2));
  * covariance against the explicit matrix product R diag(sx^2, sy^2) R^T + 0.1 I
    (numpy), the inverse against numpy.linalg.inv, the 5-sigma radius against
    ceil(5 sqrt(largest eigenvalue)) from numpy.linalg.eigvalsh, area = |sx sy|;
  * the Tukey guard-band window: zero at the outer edge, one on the image and on the last
    left-ramp sample, symmetric when both bands are equal, monotone ramps;
  * soft_window_pad: the centre is the input, each guard-band sample is the nearest edge
    value times the window (a per-sample loop), and the gradient flows only through the
    centre (the band is built without gradient);
  * the glue: identity for |h| <= 1, projection onto the unit circle (same angle) for
    |h| > 1, phase in [0, 2 pi);
  * the 3-sigma / 64-px tile renderer against the dense FP64 sum: every pixel's
    difference is at most the sum of the terms the 3-sigma cull may drop (those with
    Mahalanobis distance >= 9, i.e. G <= e^-4.5), so nothing nearer is dropped or
    counted twice.
"""
import math

import numpy as np
import torch

from oracle import model, raster

F64 = torch.float64
WL = [639e-9, 515e-9, 473e-9]
DX = 8e-6


def _raw(n, C=3, seed=0):
    rng = np.random.default_rng(seed)
    return {"means_2d": torch.as_tensor(rng.normal(0, 1, (n, 2))),
            "pre_act_scales": torch.as_tensor(rng.normal(0.5, 0.6, (n, 2))),
            "pre_act_rotation": torch.as_tensor(rng.uniform(-4, 4, n)),
            "colours": torch.as_tensor(rng.uniform(0, 1, (n, C))),
            "pre_act_phase": torch.as_tensor(rng.uniform(-10, 10, (n, C))),
            "pre_act_opacities": torch.as_tensor(rng.normal(0, 1, n)),
            "curvature": torch.as_tensor(rng.normal(0, 2, n))}


def test_activations_closed_forms():
    raw = _raw(50)
    W, H = 96, 64
    sc, rot, ph, op, mu, cv = model.apply_activations(raw, (W, H), WL, DX)
    m = raw["means_2d"].numpy()
    np.testing.assert_allclose(mu[:, 0].numpy(), (np.tanh(m[:, 0]) + 1) / 2 * W, rtol=1e-14)
    np.testing.assert_allclose(mu[:, 1].numpy(), (np.tanh(m[:, 1]) + 1) / 2 * H, rtol=1e-14)
    np.testing.assert_allclose(sc.numpy(), np.exp(raw["pre_act_scales"].numpy()) + 0.1, rtol=1e-14)
    assert float(ph.min()) >= 0 and float(ph.max()) < 2 * math.pi
    np.testing.assert_allclose(np.cos(ph.numpy()), np.cos(raw["pre_act_phase"].numpy()), atol=1e-12)
    np.testing.assert_allclose(op.numpy(), 1 / (1 + np.exp(-raw["pre_act_opacities"].numpy())),
                               rtol=1e-14)
    bound = 0.9 * min(WL) / (6 * sc.max(1).values.numpy() * DX ** 2)
    assert np.all(np.abs(cv.numpy()) < bound)
    np.testing.assert_allclose(cv.numpy(), bound * np.tanh(raw["curvature"].numpy()), rtol=1e-10)
    # gray twin binds the bound to its only wavelength; "off" gives zero curvature
    _, _, _, _, _, cg = model.apply_activations(raw, (W, H), [532e-9], DX, grayscale=True)
    np.testing.assert_allclose(cg.numpy() * min(WL), cv.numpy() * 532e-9, rtol=1e-12)
    _, _, _, _, _, c0 = model.apply_activations(raw, (W, H), WL, DX, curv_mode="off")
    assert float(c0.abs().max()) == 0.0


def test_covariance_inverse_radii_against_linear_algebra():
    raw = _raw(40, seed=1)
    ri = model.raster_inputs(raw, (80, 60), WL, DX)
    sc = ri["scales"].numpy()
    rot = raw["pre_act_rotation"].numpy()
    for g in range(40):
        c, s = math.cos(rot[g]), math.sin(rot[g])
        R = np.array([[c, -s], [s, c]])
        cov = R @ np.diag(sc[g] ** 2) @ R.T + 0.1 * np.eye(2)
        got = np.array([[ri["cov"][0][g], ri["cov"][1][g]], [ri["cov"][1][g], ri["cov"][2][g]]])
        np.testing.assert_allclose(got, cov, rtol=1e-12, atol=1e-12)
        inv = np.linalg.inv(cov)
        i00, i01, i11 = ri["inv_cov"][g].numpy()
        np.testing.assert_allclose([[i00, i01], [i01, i11]], inv, rtol=1e-10)
        lam = np.linalg.eigvalsh(cov).max()
        assert int(ri["radii"][g]) == math.ceil(5 * math.sqrt(lam))
        assert abs(float(ri["area"][g]) - abs(sc[g, 0] * sc[g, 1])) <= 1e-12
    # rotation 0: a diagonal covariance
    raw0 = dict(raw, pre_act_rotation=torch.zeros(40, dtype=F64))
    ri0 = model.raster_inputs(raw0, (80, 60), WL, DX)
    assert float(ri0["cov"][1].abs().max()) == 0.0
    np.testing.assert_allclose(ri0["cov"][0].numpy(), sc[:, 0] ** 2 + 0.1, rtol=1e-14)


def test_tukey_window():
    for total, start, size in ((40, 12, 16), (41, 10, 16), (64, 0, 64), (30, 9, 12)):
        w = model.taper_1d(total, start, size).numpy()
        np.testing.assert_array_equal(w[start:start + size], 1.0)
        if start > 0:
            assert abs(w[start - 1] - 1.0) <= 1e-15        # t = pad - 1: 0.5 (1 - cos pi)
            assert w[0] > 0 and np.all(np.diff(w[:start]) > 0)
        if start + size < total:
            assert abs(w[-1]) <= 1e-15                      # t = n - 1: 0.5 (1 + cos pi)
            assert np.all(np.diff(w[start + size:]) < 0)
        if total - start - size == start:
            # equal bands: right[t] = 0.5 (1 + cos(pi (t + 1) / p)) = left[p - 2 - t]
            np.testing.assert_allclose(w[start + size:][:-1], w[:start][:-1][::-1], atol=1e-15)


def test_soft_window_pad_by_loop():
    rng = np.random.default_rng(5)
    C, H, W, pH, pW = 2, 5, 7, 13, 16
    h = torch.as_tensor(rng.normal(size=(C, H, W)) + 1j * rng.normal(size=(C, H, W)))
    out = model.soft_window_pad(h, (pH, pW)).numpy()
    ph, pw = (pH - H) // 2, (pW - W) // 2
    wy = model.taper_1d(pH, ph, H).numpy()
    wx = model.taper_1d(pW, pw, W).numpy()
    hn = h.numpy()
    for c in range(C):
        for y in range(pH):
            for x in range(pW):
                yy = min(max(y - ph, 0), H - 1)
                xx = min(max(x - pw, 0), W - 1)
                assert abs(out[c, y, x] - hn[c, yy, xx] * wy[y] * wx[x]) <= 1e-13
    np.testing.assert_array_equal(out[:, ph:ph + H, pw:pw + W], hn)
    # gradient only through the centre
    hg = h.clone().requires_grad_()
    model.soft_window_pad(hg, (pH, pW)).abs().pow(2).sum().backward()
    np.testing.assert_allclose(hg.grad.numpy(), 2 * hn, rtol=1e-12)


def test_glue():
    rng = np.random.default_rng(6)
    h = torch.as_tensor(rng.normal(0, 0.8, (3, 20)) + 1j * rng.normal(0, 0.8, (3, 20)))
    u, ph, amp = model.glue(h)
    hn, un = h.numpy(), u.numpy()
    small = np.abs(hn) <= 1
    assert small.any() and (~small).any()
    np.testing.assert_allclose(un[small], hn[small], rtol=1e-12)
    np.testing.assert_allclose(np.abs(un[~small]), 1.0, rtol=1e-12)
    np.testing.assert_allclose(un[~small], hn[~small] / np.abs(hn[~small]), rtol=1e-12)
    assert float(ph.min()) >= 0 and float(ph.max()) < 2 * math.pi


def test_tile_renderer_within_its_cull_bound_of_the_dense_sum():
    raw = _raw(150, seed=7)
    raw["pre_act_scales"] = raw["pre_act_scales"].clamp(-1.5, 1.8)
    W, H = 140, 100                    # ragged 64-px tiles on both axes
    ri = model.raster_inputs(raw, (W, H), WL, DX)
    got = model.render_tile_torch(ri, W, H)
    args = (ri["means"], ri["inv_cov"], ri["colours"], ri["phase"], ri["opac"], ri["curv"],
            ri["area"])
    dense = raster.render_dense(*args, ri["k"], ri["dxy2"], W, H)
    diff = (got - dense).abs()
    # bound: per pixel, the colour-weighted sum of the terms with mahal >= 9
    ys, xs = torch.meshgrid(torch.arange(H, dtype=F64), torch.arange(W, dtype=F64), indexing="ij")
    dx = ri["means"][:, 0, None, None] - xs
    dy = ri["means"][:, 1, None, None] - ys
    i00, i01, i11 = (ri["inv_cov"][:, j, None, None] for j in range(3))
    mahal = dx * dx * i00 + 2 * dx * dy * i01 + dy * dy * i11
    far = (mahal >= 9.0).to(F64)
    term = ri["opac"][:, None, None] * torch.exp(torch.clamp(-0.5 * mahal, min=-50)) * far
    bound = torch.einsum("ghw,gc->chw", term, ri["colours"])
    assert bool((diff <= bound + 1e-12).all())
    assert float(diff.max()) > 0                      # some far terms are dropped
    rel = float(diff.norm() / dense.abs().norm())
    assert rel < 1e-2
