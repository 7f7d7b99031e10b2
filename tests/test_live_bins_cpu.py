"""CPU check of the live binning's row-interval test (k_count_live in csrc/raster.hip).

The live count keeps a tile of a Gaussian's rectangle if the tile meets, in its row of
tiles, the interval of pixel x at which the maximum of the log2-scaled exponent q over the
row's y range reaches QSKIP - 1/16, widened by 0.25 px + 1e-3 of the extent.  It must keep
every key the raster kernels evaluate: those whose closed-form maximum of q over the tile
(tile_qmax) reaches QSKIP = -24.  Checked here in float32 on seeded Gaussians of every
shape the activations produce (isotropic to strongly elongated, any rotation, centres
inside and outside the frame, partial tiles at the right and bottom edges), against a
brute-force maximum over each tile's pixels too, and the test must prune: fewer keys than
the full rectangles.
"""
import numpy as np

F = np.float32
QSCALE = F(-0.5 * 1.4426950408889634)
QSKIP, MARGIN = F(-24.0), F(0.0625)


def tile_qmax(mx, my, a00, a01, a11, x0, x1, y0, y1):
    """csrc/raster.hip tile_qmax in float32."""
    dxlo, dxhi, dylo, dyhi = mx - x1, mx - x0, my - y1, my - y0
    inside = (dxlo <= 0) & (dxhi >= 0) & (dylo <= 0) & (dyhi >= 0)

    def qf(dx, dy):
        return dx * (a00 * dx + F(2) * a01 * dy) + a11 * dy * dy
    ry, rx = -a01 / a11, -a01 / a00
    b = qf(dxlo, np.clip(ry * dxlo, dylo, dyhi))
    b = np.maximum(b, qf(dxhi, np.clip(ry * dxhi, dylo, dyhi)))
    b = np.maximum(b, qf(np.clip(rx * dylo, dxlo, dxhi), dylo))
    b = np.maximum(b, qf(np.clip(rx * dyhi, dxlo, dxhi), dyhi))
    return np.where(inside, F(0), b)


def row_interval(mx, my, a00, a01, a11, y0, y1):
    """The live count's interval [xa, xb] of one row of tiles (empty: xa > xb)."""
    thr = QSKIP - MARGIN
    beta = a00 * a11 - a01 * a01
    rx, ry, ia = -a01 / a00, -a01 / a11, F(1) / np.abs(a00)
    dxr = np.sqrt(np.maximum(thr * a11 / beta, F(0)))
    dyr = ry * dxr
    dymax = np.sqrt(np.maximum(a00 * thr / beta, F(0)))
    c0 = a00 * thr
    dylo, dyhi = my - y1, my - y0
    hit = (dyhi >= -dymax) & (dylo <= dymax)
    er, el = np.clip(dyr, dylo, dyhi), np.clip(-dyr, dylo, dyhi)
    dr = np.where((dyr >= dylo) & (dyr <= dyhi), dxr,
                  rx * er + np.sqrt(np.maximum(c0 - beta * er * er, F(0))) * ia)
    dl = np.where((-dyr >= dylo) & (-dyr <= dyhi), -dxr,
                  rx * el - np.sqrt(np.maximum(c0 - beta * el * el, F(0))) * ia)
    xa = mx - (dr + (F(0.25) + F(1e-3) * np.abs(dr)))
    xb = mx - (dl - (F(0.25) + F(1e-3) * np.abs(dl)))
    return np.where(hit, xa, F(1)), np.where(hit, xb, F(0))


def make_keys(seed, n, W, H):
    """Gaussians as the prologue makes them (covariance R diag(s^2) R^T + 0.1 I, inverse,
    log2 scaling) and every (Gaussian, tile) key of their 5-sigma rectangles."""
    rng = np.random.default_rng(seed)
    mx = rng.uniform(-20, W + 20, n)
    my = rng.uniform(-20, H + 20, n)
    s = np.exp(rng.uniform(-2.5, 3.5, (n, 2))) + 0.1
    th = rng.uniform(0, 2 * np.pi, n)
    c, sn = np.cos(th), np.sin(th)
    c00 = s[:, 0] ** 2 * c * c + s[:, 1] ** 2 * sn * sn + 0.1
    c01 = (s[:, 0] ** 2 - s[:, 1] ** 2) * c * sn
    c11 = s[:, 0] ** 2 * sn * sn + s[:, 1] ** 2 * c * c + 0.1
    det = c00 * c11 - c01 * c01
    i00, i01, i11 = c11 / det, -c01 / det, c00 / det
    mid = 0.5 * (c00 + c11)
    rad = np.ceil(5 * np.sqrt(mid + np.sqrt(((c00 - c11) / 2) ** 2 + c01 ** 2))).astype(int)
    gx, gy = (W + 15) // 16, (H + 15) // 16
    lx = np.clip(((mx - rad) / 16).astype(int), 0, gx)
    hx = np.clip(((mx + rad + 15) / 16).astype(int), 0, gx)
    ly = np.clip(((my - rad) / 16).astype(int), 0, gy)
    hy = np.clip(((my + rad + 15) / 16).astype(int), 0, gy)
    g, tx, ty = [], [], []
    for k in range(n):
        for yy in range(ly[k], hy[k]):
            for xx in range(lx[k], hx[k]):
                g.append(k)
                tx.append(xx)
                ty.append(yy)
    g, tx, ty = np.array(g), np.array(tx), np.array(ty)
    a = [F(QSCALE) * v.astype(F) for v in (i00, i01, i11)]
    return dict(mx=mx.astype(F)[g], my=my.astype(F)[g], a00=a[0][g], a01=a[1][g],
                a11=a[2][g], tx=tx, ty=ty)


def test_row_intervals_keep_every_evaluated_key():
    W, H = 328, 248  # partial tiles on the right (8 px) and at the bottom (8 px)
    kept_total = live_total = keys_total = 0
    for seed in range(4):
        k = make_keys(seed, 1500, W, H)
        x0 = (k["tx"] * 16).astype(F)
        x1 = (np.minimum(k["tx"] * 16 + 16, W) - 1).astype(F)
        y0 = (k["ty"] * 16).astype(F)
        y1 = (np.minimum(k["ty"] * 16 + 16, H) - 1).astype(F)
        args = (k["mx"], k["my"], k["a00"], k["a01"], k["a11"])
        kept = tile_qmax(*args, x0, x1, y0, y1) >= QSKIP
        xa, xb = row_interval(*args, y0, y1)
        live = (x1 >= xa) & (x0 <= xb)
        assert not np.any(kept & ~live), f"seed {seed}: {int((kept & ~live).sum())} keys lost"
        # brute force over the tile's pixels: a key with a pixel at q >= QSKIP is kept
        yy, xx = np.mgrid[0:16, 0:16]
        px = x0[:, None] + xx.ravel()[None, :]
        py = y0[:, None] + yy.ravel()[None, :]
        dx, dy = args[0][:, None] - px, args[1][:, None] - py
        q = dx * (args[2][:, None] * dx + F(2) * args[3][:, None] * dy) + args[4][:, None] * dy * dy
        q = np.where((px <= x1[:, None]) & (py <= y1[:, None]), q, -np.inf)
        assert not np.any((q.max(1) >= QSKIP) & ~live)
        kept_total += int(kept.sum())
        live_total += int(live.sum())
        keys_total += kept.size
    # the interval test prunes (close to the kernels' own set)
    assert kept_total <= live_total < 0.9 * keys_total
    assert live_total < 1.05 * kept_total
