"""Pins oracle.optics (utils/propagator.py:10-64, 202-290 and the odak BLASM kernel it calls).

The checks use the mathematics of the angular spectrum method rather than the
oracle's own expressions:
  * z = 0: H = 1 on the whole grid, so pad -> FFT -> IFFT -> crop returns the input field;
  * |H| = 1 inside the band and 0 outside (the band limit is inactive at the small-grid
    sizes below), so with an open aperture the propagated energy on the 2x grid equals
    the input energy (Parseval);
  * a plane wave on one exact FFT bin (p, q) is an eigenfunction of the propagation: it
    comes out multiplied by exp(i 2 pi z sqrt(1/lambda^2 - (p/X)^2 - (q/Y)^2)), written
    here from the formula with numpy's fftfreq, independent of the oracle's centred grid
    and its fftshift bookkeeping;
  * propagating by z1 then z2 equals propagating by z1 + z2 (H(z1) H(z2) = H(z1 + z2));
  * the whole field_at pipeline against a numpy.fft restatement that never shifts
    (transfer function evaluated directly in FFT order with fftfreq);
  * zero_pad2 / crop_center invert each other (even sizes, the sizes the build uses),
    and the centre crop of an explicit size against hand-worked indices;
  * the band limit against a brute-force scan of every bin (with odak's axis naming,
    which bounds the column frequencies with the row extent; see the test);
  * the circular aperture against a per-sample loop.
odak itself is not installed, so the odak half-sample grid convention stays "parity
unpinned"; only its symmetry and its distance from the exact bins are checked.
"""
import math

import numpy as np
import torch

from oracle import optics

DX = 8e-6
LAM = 515e-9


def _prop(pH, pW, zs, wl=(LAM,), grid="fft", aperture=1e9):
    return optics.BLASM((pH, pW), list(wl), DX, list(zs), aperture_size=aperture, grid=grid)


def _field(C, n, m, seed):
    rng = np.random.default_rng(seed)
    return torch.as_tensor(rng.normal(size=(C, n, m)) + 1j * rng.normal(size=(C, n, m)))


def test_zero_distance_is_identity():
    u = _field(2, 24, 40, 0)
    for grid in ("fft", "odak"):
        prop = _prop(24, 40, [0.0], wl=(639e-9, 473e-9), grid=grid)
        for c in range(2):
            out = prop.field_at(u, 0, c)
            np.testing.assert_allclose(out.numpy(), u[c].numpy(), rtol=0, atol=1e-12)


def test_energy_conserved_with_open_aperture():
    """Full 2x grid (no crop): sum |ifft(H fft(U))|^2 = sum |U|^2 when |H| = 1 everywhere."""
    prop = _prop(32, 48, [1e-3, 5e-3])
    for d in range(2):
        H = prop.H[d, 0]
        assert np.allclose(np.abs(H.numpy()), 1.0)          # band limit inactive here
        U = optics.zero_pad2(_field(1, 32, 48, 1 + d)[0])
        S = torch.fft.fftshift(torch.fft.fft2(U))
        out = torch.fft.ifft2(torch.fft.ifftshift(H * S))
        e_in, e_out = float((U.abs() ** 2).sum()), float((out.abs() ** 2).sum())
        assert abs(e_out - e_in) <= 1e-12 * e_in


def _full_grid_propagate(prop, U, d, c):
    S = torch.fft.fftshift(torch.fft.fft2(U))
    return torch.fft.ifft2(torch.fft.ifftshift(prop.H[d, c] * S))


def test_plane_wave_on_an_fft_bin_picks_up_the_transfer_function():
    nu, nv = 64, 96
    z = 3e-3
    prop = _prop(nu // 2, nv // 2, [z])
    yy, xx = np.mgrid[0:nu, 0:nv]
    fr = np.fft.fftfreq(nu, DX)     # cycles per metre of each bin, FFT order
    fc = np.fft.fftfreq(nv, DX)
    for p, q in [(0, 0), (3, -5), (-20, 17), (31, 47), (-32, -48)]:
        U = torch.as_tensor(np.exp(2j * np.pi * (p * yy / nu + q * xx / nv)))
        out = _full_grid_propagate(prop, U, 0, 0)
        f2 = fr[p % nu] ** 2 + fc[q % nv] ** 2
        expect = np.exp(2j * np.pi * z * np.sqrt(1 / LAM ** 2 - f2))
        np.testing.assert_allclose(out.numpy(), expect * U.numpy(), rtol=0, atol=1e-10)


def test_propagation_composes():
    z1, z2 = 1.3e-3, 2.9e-3
    prop = _prop(24, 32, [z1, z2, z1 + z2], wl=(639e-9,))
    U = optics.zero_pad2(_field(1, 24, 32, 5)[0])
    two = _full_grid_propagate(prop, _full_grid_propagate(prop, U, 0, 0), 1, 0)
    one = _full_grid_propagate(prop, U, 2, 0)
    np.testing.assert_allclose(two.numpy(), one.numpy(), rtol=0, atol=1e-10)


def test_field_at_matches_unshifted_numpy_restatement():
    pH, pW, z, wl = 20, 36, 4e-3, 473e-9
    prop = _prop(pH, pW, [z], wl=(wl,))
    u = _field(1, pH, pW, 7)[0].numpy()
    nu, nv = 2 * pH, 2 * pW
    U = np.zeros((nu, nv), complex)
    U[pH - pH // 2:pH - pH // 2 + pH, pW - pW // 2:pW - pW // 2 + pW] = u
    FY, FX = np.meshgrid(np.fft.fftfreq(nu, DX), np.fft.fftfreq(nv, DX), indexing="ij")
    Hn = np.exp(2j * np.pi * z * np.sqrt(1 / wl ** 2 - FX ** 2 - FY ** 2))
    out = np.fft.ifft2(np.fft.fft2(U) * Hn)[nu // 4:nu // 4 + nu // 2, nv // 4:nv // 4 + nv // 2]
    got = prop.field_at(torch.as_tensor(u)[None], 0, 0).numpy()
    np.testing.assert_allclose(got, out, rtol=0, atol=1e-10)


def test_pad_and_crop_conventions():
    u = _field(1, 6, 10, 9)
    p = optics.zero_pad2(u)
    assert p.shape[-2:] == (12, 20)
    np.testing.assert_array_equal(p[0, 3:9, 5:15].numpy(), u[0].numpy())
    assert float(p.abs().sum()) == float(u.abs().sum())
    np.testing.assert_array_equal(optics.crop_center(p).numpy(), u.numpy())
    # explicit size: rows 12//2 - 4//2 = 4 .. 8, columns 20//2 - 6//2 = 7 .. 13
    np.testing.assert_array_equal(optics.crop_center(p, (4, 6)).numpy(), p[..., 4:8, 7:13].numpy())


def test_band_limit_admits_exactly_the_frequencies_below_the_bound():
    """Long distance on a small non-square grid makes the band limit active.  odak names
    the column frequencies FX and bounds them with the extent X = dx * nu of the first
    (row) axis, and the row frequencies FY with Y = dx * nv (the naming of
    ``FY, FX = meshgrid(fx, fy, indexing='ij')``); on a square grid the two coincide.  A
    brute-force scan checks that exactly the bins below both bounds are admitted."""
    nu, nv, z = 64, 40, 0.2
    H = optics.transfer_function(nu, nv, DX, LAM, z, grid="fft")
    X, Y = DX * nu, DX * nv
    lim_col = 1 / (LAM * math.sqrt((2 * z / X) ** 2 + 1))
    lim_row = 1 / (LAM * math.sqrt((2 * z / Y) ** 2 + 1))
    fr = (np.arange(nu) - nu // 2) / X
    fc = (np.arange(nv) - nv // 2) / Y
    on = np.abs(H) > 0.5
    assert 0 < on.sum() < on.size
    for i in range(nu):
        for j in range(nv):
            assert on[i, j] == (abs(fr[i]) < lim_row and abs(fc[j]) < lim_col), (i, j)
    Hs = optics.transfer_function(48, 48, DX, LAM, z, grid="fft")
    np.testing.assert_array_equal(np.abs(Hs) > 0.5, (np.abs(Hs) > 0.5).T)


def test_odak_grid_is_symmetric_and_near_the_exact_bins():
    for n in (64, 2816, 4480):
        f = optics.freq_grid(n, DX, "odak")
        e = optics.freq_grid(n, DX, "fft")
        np.testing.assert_allclose(f, -f[::-1], rtol=0, atol=1e-9)
        assert np.abs(f - e).max() <= 1.0 / (DX * n)     # within one bin of the exact grid


def test_circular_mask_matches_loop():
    px, py, r = 9, 14, 4.3
    m = optics.circular_mask(px, py, r)
    for i in range(px):
        for j in range(py):
            x = -px / 2 + i * px / (px - 1)
            y = -py / 2 + j * py / (py - 1)
            assert m[i, j] == (1.0 if math.hypot(x, y) < r else 0.0)


def test_custom_aperture_scales_each_bin_by_its_square():
    """A custom aperture a [pH, pW] (utils/propagator.py:199-200: zero-padded to the 2x grid;
    custom() multiplies the shifted spectrum and H by it): a plane wave on FFT bin (p, q)
    comes out multiplied by a_pad(p, q)^2 H(f), a_pad read at the bin's centred position
    (p + nu/2 mod nu), so bins outside the padded support vanish."""
    pH, pW, z = 16, 24, 2e-3
    nu, nv = 2 * pH, 2 * pW
    rng = np.random.default_rng(11)
    a = rng.uniform(0.3, 1.2, (pH, pW))
    prop = optics.BLASM((pH, pW), [LAM], DX, [z], grid="fft", aperture=a)
    a_pad = np.zeros((nu, nv))
    a_pad[pH - pH // 2:pH - pH // 2 + pH, pW - pW // 2:pW - pW // 2 + pW] = a
    yy, xx = np.mgrid[0:pH, 0:pW]
    fr, fc = np.fft.fftfreq(nu, DX), np.fft.fftfreq(nv, DX)
    for p, q in [(0, 0), (3, -5), (-7, 11), (15, 2), (-16, -24), (9, 13)]:
        # the plane wave restricted to the hologram (what field_at pads and propagates) is
        # not a single bin of the 2x grid; use the full-grid wave and the uncropped pipeline
        U = torch.as_tensor(np.exp(2j * np.pi * (p * np.arange(nu)[:, None] / nu +
                                                 q * np.arange(nv)[None, :] / nv)))
        ap = prop.aperture.to(U.dtype)
        S = torch.fft.fftshift(torch.fft.fft2(U)) * ap
        out = torch.fft.ifft2(torch.fft.ifftshift((prop.H[0, 0] * ap) * S))
        ci, cj = (p % nu + nu // 2) % nu, (q % nv + nv // 2) % nv
        f2 = fr[p % nu] ** 2 + fc[q % nv] ** 2
        expect = a_pad[ci, cj] ** 2 * np.exp(2j * np.pi * z * np.sqrt(1 / LAM ** 2 - f2))
        np.testing.assert_allclose(out.numpy(), expect * U.numpy(), rtol=0, atol=1e-10)
    assert np.isclose(float(prop.aperture.sum()), a.sum())
