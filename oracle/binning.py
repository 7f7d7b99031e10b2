"""Tile binning of the QPF rasterizer, restated in numpy (bit-exact).

Restates
  * ``getRect``                  cuda/forward.cu:8-13
  * ``preprocessKernel``         cuda/forward.cu:15-26
  * inclusive scan + P readback  cuda/raster.cu:62-75
  * ``duplicateWithKeysKernel``  cuda/forward.cu:28-47
  * 64-bit key sort              cuda/raster.cu:85-93
  * ``identifyTileRangesKernel`` cuda/forward.cu:49-60

The rectangle arithmetic is done in float32 exactly as the CUDA code does it:
``(int)((pos - r) / 16)`` and ``(int)((((pos + r) + 16) - 1) / 16)`` with one
float32 rounding per operation, truncation toward zero, then clamping to
``[0, grid]``.

Pinned by (tests/test_oracle_binning.py):
  * a brute-force per-tile membership scan written independently of the
    emission/sort path (each tile lists exactly the Gaussians whose rectangle
    covers it, in ascending id order);
  * invariants: sum(tiles_touched) == P, ranges tile the key list, empty tiles
    stay (0, 0), P == 0 gives ``point_list == [0]``;
  * hand-worked rectangles at tile borders and negative coordinates.
"""
import numpy as np

BLOCK = 16


def tile_grid(width, height):
    """Tile grid (gx, gy) of 16x16 tiles, cuda/raster.cu:50."""
    return (width + BLOCK - 1) // BLOCK, (height + BLOCK - 1) // BLOCK


def get_rect(means, radii, width, height):
    """Return (lo_x, lo_y, hi_x, hi_y) int64 arrays; cuda/forward.cu:8-13."""
    gx, gy = tile_grid(width, height)
    m = np.ascontiguousarray(means, dtype=np.float32)
    r = np.asarray(radii).astype(np.int64)
    rf = r.astype(np.float32)
    px, py = m[:, 0], m[:, 1]
    b = np.float32(BLOCK)
    one = np.float32(1.0)
    with np.errstate(invalid="ignore", over="ignore"):
        lox = np.trunc((px - rf) / b)
        loy = np.trunc((py - rf) / b)
        hix = np.trunc((((px + rf) + b) - one) / b)
        hiy = np.trunc((((py + rf) + b) - one) / b)

    def clamp(v, g):
        v = np.nan_to_num(v.astype(np.float64), nan=0.0, posinf=g, neginf=0.0)
        return np.minimum(g, np.maximum(0, v)).astype(np.int64)

    return clamp(lox, gx), clamp(loy, gy), clamp(hix, gx), clamp(hiy, gy)


def tiles_touched(means, radii, width, height):
    """Per-Gaussian count of covered tiles; cuda/forward.cu:15-26."""
    lox, loy, hix, hiy = get_rect(means, radii, width, height)
    r = np.asarray(radii).astype(np.int64)
    ok = (r > 0) & (hix > lox) & (hiy > loy)
    return np.where(ok, (hiy - loy) * (hix - lox), 0).astype(np.int64)


def bin_gaussians(means, radii, width, height):
    """Full binning: returns a dict with
    ``tiles_touched`` [N] int64, ``offsets`` [N] (inclusive scan), ``P``,
    ``keys`` [P] emission-order tile ids, ``vals`` [P] emission-order ids,
    ``point_list`` [max(P,1)] int32 (sorted ids; ``[0]`` when P == 0),
    ``ranges`` [T, 2] int32.
    """
    gx, gy = tile_grid(width, height)
    T = gx * gy
    n = np.asarray(radii).shape[0]
    cnt = tiles_touched(means, radii, width, height)
    offsets = np.cumsum(cnt)
    P = int(offsets[-1]) if n else 0
    ranges = np.zeros((T, 2), dtype=np.int32)
    if P <= 0:
        return dict(tiles_touched=cnt, offsets=offsets, P=0,
                    keys=np.zeros(0, np.int64), vals=np.zeros(0, np.int64),
                    point_list=np.zeros(1, np.int32), ranges=ranges)
    lox, loy, hix, hiy = get_rect(means, radii, width, height)
    keys = np.empty(P, np.int64)
    vals = np.empty(P, np.int64)
    # vectorised emission: for every Gaussian, its rectangle in row-major order
    g_of = np.repeat(np.arange(n), cnt)
    start = np.concatenate([[0], offsets[:-1]])
    local = np.arange(P) - start[g_of]
    w = (hix - lox)[g_of]
    ty = loy[g_of] + local // np.maximum(w, 1)
    tx = lox[g_of] + local % np.maximum(w, 1)
    keys[:] = ty * gx + tx
    vals[:] = g_of
    # reference sorts the 64-bit key (tile << 32) | id on all 64 bits
    key64 = (keys.astype(np.uint64) << np.uint64(32)) | vals.astype(np.uint64)
    order = np.argsort(key64, kind="stable")
    sk = keys[order]
    point_list = vals[order].astype(np.int32)
    # ranges: [first, last + 1) per tile present; absent tiles stay (0, 0)
    change = np.flatnonzero(np.diff(sk)) + 1
    starts = np.concatenate([[0], change])
    ends = np.concatenate([change, [P]])
    ranges[sk[starts], 0] = starts
    ranges[sk[starts], 1] = ends
    return dict(tiles_touched=cnt, offsets=offsets, P=P, keys=keys, vals=vals,
                point_list=point_list, ranges=ranges)


def brute_force_tile_lists(means, radii, width, height):
    """Independent check: for every tile, the ascending ids whose rect covers it."""
    gx, gy = tile_grid(width, height)
    lox, loy, hix, hiy = get_rect(means, radii, width, height)
    r = np.asarray(radii).astype(np.int64)
    out = []
    for t in range(gx * gy):
        ty, tx = divmod(t, gx)
        sel = (r > 0) & (lox <= tx) & (tx < hix) & (loy <= ty) & (ty < hiy)
        out.append(np.flatnonzero(sel))
    return out


def sub_bins(bins, tiles):
    """The key list restricted to the given tiles (ascending tile ids): the same per-tile
    Gaussian lists, re-based so that they cover [0, P') in tile order.  Lets a caller
    evaluate a large frame a few tile rows at a time."""
    rng = bins["ranges"]
    pl = bins["point_list"]
    parts = [pl[rng[t, 0]:rng[t, 1]] for t in tiles]
    out_pl = np.concatenate(parts).astype(np.int32) if parts else np.zeros(0, np.int32)
    ranges = np.zeros_like(rng)
    off = 0
    for t, p in zip(tiles, parts):
        if len(p):
            ranges[t] = (off, off + len(p))
            off += len(p)
    return dict(P=off, point_list=out_pl if off else np.zeros(1, np.int32), ranges=ranges)
