"""Pins oracle.raster (cuda/rasterizer.py:74-90, cuda/forward.cu:62-162, cuda/backward.cu:11-168)."""
import numpy as np
import torch

from oracle import binning, raster
from oracle.step import render_binned_diff


def make_inputs(n, W, H, C, seed, curv_scale=0.02, smin=1.0, smax=6.0):
    rng = np.random.default_rng(seed)
    means = np.stack([rng.uniform(0, W, n), rng.uniform(0, H, n)], 1)
    s = np.exp(rng.uniform(np.log(smin), np.log(smax), (n, 2)))
    rot = rng.uniform(0, np.pi, n)
    c, si = np.cos(rot), np.sin(rot)
    c00 = s[:, 0] ** 2 * c * c + s[:, 1] ** 2 * si * si + 0.1
    c01 = (s[:, 0] ** 2 - s[:, 1] ** 2) * c * si
    c11 = s[:, 0] ** 2 * si * si + s[:, 1] ** 2 * c * c + 0.1
    det = c00 * c11 - c01 * c01
    inv = np.stack([c11 / det, -c01 / det, c00 / det], 1)
    lam = np.array([639e-9, 515e-9, 473e-9])[:C] if C == 3 else np.array([515e-9])
    k = 2 * np.pi / lam
    dxy2 = (8e-6) ** 2
    kap = rng.normal(0, curv_scale, n)
    curv = 2 * kap / (k.max() * dxy2)
    mid = 0.5 * (c00 + c11)
    disc = np.sqrt(((c00 - c11) * 0.5) ** 2 + c01 ** 2)
    radii = np.ceil(5 * np.sqrt(mid + disc)).astype(np.int32)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return dict(means=f32(means), inv_cov=f32(inv), colours=f32(rng.uniform(0, 1, (n, C))),
                phase=f32(rng.uniform(0, 2 * np.pi, (n, C))), opac=f32(rng.uniform(0.2, 1, n)),
                curv=f32(curv), area=f32(np.abs(s[:, 0] * s[:, 1])), radii=radii, k=f32(k),
                dxy2=dxy2)


def _args(d):
    return (d["means"], d["inv_cov"], d["colours"], d["phase"], d["opac"], d["curv"], d["area"])


def test_closed_form_isotropic_flat_and_chirp():
    W, H = 40, 30
    s, a, o, phi = 3.0, 0.7, 0.6, 1.1
    mx, my = 17.3, 12.6
    k = np.array([2 * np.pi / 515e-9])
    dxy2 = 64e-12
    for c_val in (0.0, 2e6):
        area = s * s
        out = raster.render_dense(np.array([[mx, my]]), np.array([[1 / s ** 2, 0, 1 / s ** 2]]),
                                  np.array([[a]]), np.array([[phi]]), np.array([o]),
                                  np.array([c_val]), np.array([area]), k, dxy2, W, H)
        yy, xx = np.mgrid[0:H, 0:W]
        r2 = (xx - mx) ** 2 + (yy - my) ** 2
        m = r2 / s ** 2
        expect = a * o * np.exp(-np.minimum(m / 2, 50)) * np.exp(
            1j * (phi + k[0] * 0.5 * c_val * m * area * dxy2))
        np.testing.assert_allclose(out[0].numpy(), expect, rtol=1e-9, atol=1e-13)


def test_superposition():
    d = make_inputs(40, 48, 40, 3, 1)
    full = raster.render_dense(*_args(d), d["k"], d["dxy2"], 48, 40)
    idx = np.arange(40)
    parts = 0
    for sel in (idx[:17], idx[17:]):
        parts = parts + raster.render_dense(*(a[sel] for a in _args(d)), d["k"], d["dxy2"], 48, 40)
    np.testing.assert_allclose(full.numpy(), parts.numpy(), rtol=1e-12, atol=1e-14)


def test_binned_close_to_dense():
    d = make_inputs(120, 80, 64, 3, 2)
    dense = raster.render_dense(*_args(d), d["k"], d["dxy2"], 80, 64)
    binned, bins = raster.render_binned(*_args(d), d["radii"], d["k"], d["dxy2"], 80, 64)
    rel = float((binned - dense).abs().norm() / dense.abs().norm())
    assert rel < 1e-5, rel
    assert bins["P"] > 0


def test_closed_form_adjoint_matches_autograd_and_fd():
    W, H, C = 40, 36, 3
    d = make_inputs(25, W, H, C, 3, curv_scale=0.05)
    bins = binning.bin_gaussians(d["means"], d["radii"], W, H)
    rng = np.random.default_rng(9)
    grad = torch.complex(torch.as_tensor(rng.normal(size=(C, H, W))),
                         torch.as_tensor(rng.normal(size=(C, H, W))))
    closed = raster.render_backward(grad, *_args(d), d["k"], d["dxy2"], W, H, bins)
    names = ["means", "inv_cov", "colours", "phase", "opac", "curv", "area"]
    leaves = {n: torch.as_tensor(d[n], dtype=torch.float64).requires_grad_() for n in names}
    ri = dict(means=leaves["means"], inv_cov=leaves["inv_cov"], colours=leaves["colours"],
              phase=leaves["phase"], opac=leaves["opac"], curv=leaves["curv"], area=leaves["area"],
              k=torch.as_tensor(d["k"], dtype=torch.float64), dxy2=d["dxy2"])
    U = render_binned_diff(ri, W, H, bins)
    L = (U.real * grad.real).sum() + (U.imag * grad.imag).sum()
    L.backward()
    for n, cl in zip(names, closed):
        ag = leaves[n].grad
        np.testing.assert_allclose(cl.numpy(), ag.numpy().reshape(cl.shape), rtol=1e-9,
                                   atol=1e-9 * float(ag.abs().max()) + 1e-300, err_msg=n)
    # central finite differences of the dense forward on a few entries (h = 1e-4 relative)
    def Lval(dd):
        u = raster.render_dense(*(dd[n] for n in names), d["k"], d["dxy2"], W, H)
        return float((u.real * grad.real).sum() + (u.imag * grad.imag).sum())
    base = {n: np.asarray(d[n], np.float64).copy() for n in names}
    for n, cl in zip(names, closed):
        arr = base[n]
        flat = arr.reshape(-1)
        for j in (0, flat.size // 2, flat.size - 1):
            h = 1e-4 * max(1.0, abs(flat[j]))
            if n == "curv":
                h = 1e-4 * max(abs(flat[j]), 1e3)
            plus = {m: base[m].copy() for m in names}
            minus = {m: base[m].copy() for m in names}
            plus[n].reshape(-1)[j] += h
            minus[n].reshape(-1)[j] -= h
            fd = (Lval(plus) - Lval(minus)) / (2 * h)
            an = float(cl.reshape(-1)[j])
            assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)) + 1e-6, (n, j, fd, an)


def test_adjoint_dot_product():
    W, H, C = 32, 32, 1
    d = make_inputs(15, W, H, C, 4)
    bins = binning.bin_gaussians(d["means"], d["radii"], W, H)
    rng = np.random.default_rng(5)
    w = torch.complex(torch.as_tensor(rng.normal(size=(C, H, W))), torch.as_tensor(rng.normal(size=(C, H, W))))
    v = rng.normal(size=(15, C))
    # J v along the colour direction: U is linear in colours, so J v = U(colours = v)
    dd = dict(d)
    dd["colours"] = v
    Jv, _ = raster.render_binned(*_args(dd), d["radii"], d["k"], d["dxy2"], W, H, bins)
    lhs = float((Jv.real * w.real).sum() + (Jv.imag * w.imag).sum())
    JTw = raster.render_backward(w, *_args(d), d["k"], d["dxy2"], W, H, bins)[2]
    rhs = float((JTw.numpy() * v).sum())
    assert abs(lhs - rhs) < 1e-10 * max(1, abs(lhs))


def test_window_render_equals_binned_crop():
    """render_binned_window on a window not aligned to tiles equals the crop of the whole
    binned render (same pairs, FP64)."""
    d = make_inputs(300, 80, 70, 3, 4)
    full, bins = raster.render_binned(*_args(d), d["radii"], d["k"], d["dxy2"], 80, 70)
    win = raster.render_binned_window(*_args(d), d["k"], d["dxy2"], 80, 70, bins, 21, 9, 50, 61,
                                      chunk_tiles=3)
    np.testing.assert_allclose(win.numpy(), full[:, 9:70, 21:71].numpy(), rtol=1e-12, atol=1e-14)
