"""Pins of oracle.export (DPAC encoding and PSNR / SSIM) by hand-worked cases, closed
forms and brute force (export_poh.py:110-119,173-178; train_2d_gaussian_rgb.py:98-112)."""
import numpy as np

from oracle import export as oexp


def test_dpac_checkerboard_hand_worked():
    ph = np.full((2, 2), 1.0)
    amp = np.full((2, 2), np.cos(0.3))
    got = oexp.dpac(ph, amp)
    want = np.array([[0.7, 1.3], [1.3, 0.7]])     # (even, even), (odd, odd): phase - offset
    assert np.allclose(got, want, atol=1e-12)


def test_dpac_wraps_into_zero_two_pi():
    ph = np.array([[0.1, 6.2]])
    amp = np.array([[0.0, 0.0]])                  # offset pi / 2
    got = oexp.dpac(ph, amp)
    assert np.allclose(got, [[0.1 - np.pi / 2 + 2 * np.pi, 6.2 + np.pi / 2 - 2 * np.pi]])
    assert (got >= 0).all() and (got < 2 * np.pi).all()


def test_quantize_half_to_even_and_wrap():
    assert oexp.quantize(np.array([0.0, 2.5, 3.5, 255.4, 255.6, 2.49])).tolist() == [0, 2, 4, 255, 0, 2]


def test_levels_unit_amplitude():
    # amplitude 1: offset 0, level = round(phase / 2pi * 256)
    ph = np.array([[0.0, 2.2, 3.7, 255.4, 255.6]]) / 256.0 * 2 * np.pi
    lv, v = oexp.dpac_levels(np.exp(1j * ph))
    assert np.allclose(v, [[0.0, 2.2, 3.7, 255.4, 255.6]], atol=1e-9)
    assert lv.tolist() == [[0, 2, 4, 255, 0]]


def test_levels_amplitude_clamped_and_zero_field():
    f = np.array([[3.0 + 0j, 0.0 + 0j]])           # |h| > 1 clamps to 1; h = 0 has phase 0
    lv, v = oexp.dpac_levels(f)
    assert lv[0, 0] == 0                           # phase 0, offset 0
    assert np.isclose(v[0, 1], 256.0 * (np.pi / 2) / (2 * np.pi))   # odd pixel: 0 + pi/2


def test_psnr_closed_form():
    t = np.zeros((3, 8, 9))
    r = np.full((3, 8, 9), 0.1)
    assert np.isclose(oexp.psnr(t, r), 20.0)


def test_ssim_identity_and_bruteforce():
    rng = np.random.default_rng(0)
    x = rng.uniform(size=(13, 17))
    y = np.clip(x + 0.1 * rng.normal(size=x.shape), 0, 1)
    assert np.isclose(oexp.ssim_channel(x, x), 1.0, atol=1e-12)
    assert np.isclose(oexp.ssim_channel(x, y), oexp.ssim_channel_bruteforce(x, y), atol=1e-12)


def test_crop_center_and_evaluate_shapes():
    rec = np.arange(2 * 1 * 8 * 10, dtype=np.float64).reshape(2, 1, 8, 10) / 200.0
    c = oexp.crop_center(rec, 4, 6)
    assert c.shape == (2, 1, 4, 6) and c[0, 0, 0, 0] == rec[0, 0, 2, 2]
    tgt = np.clip(oexp.crop_center(rec, 4, 6) + 0.01, 0, 1)
    out = oexp.evaluate(rec, tgt, 4, 6)
    assert len(out["psnr"]) == 2 and np.isclose(out["psnr"][0], 40.0)
