"""Pins oracle.binning (the restatement of cuda/forward.cu:8-60 and cuda/raster.cu:30-105)."""
import numpy as np

from oracle import binning


def _scene(n, W, H, seed, rmax=40):
    rng = np.random.default_rng(seed)
    means = np.stack([rng.uniform(-20, W + 20, n), rng.uniform(-20, H + 20, n)], 1).astype(np.float32)
    radii = rng.integers(-2, rmax, n).astype(np.int32)
    return means, radii


def test_matches_brute_force_membership():
    for seed, (W, H) in enumerate([(64, 48), (100, 37), (256, 256), (17, 200)]):
        means, radii = _scene(300, W, H, seed)
        b = binning.bin_gaussians(means, radii, W, H)
        lists = binning.brute_force_tile_lists(means, radii, W, H)
        gx, gy = binning.tile_grid(W, H)
        assert b["ranges"].shape == (gx * gy, 2)
        for t, ids in enumerate(lists):
            s, e = b["ranges"][t]
            if ids.size == 0:
                assert (s, e) == (0, 0)
            else:
                np.testing.assert_array_equal(b["point_list"][s:e], ids)
        assert b["P"] == int(b["tiles_touched"].sum()) == sum(len(x) for x in lists)


def test_ranges_tile_the_key_list():
    means, radii = _scene(500, 128, 96, 7)
    b = binning.bin_gaussians(means, radii, 128, 96)
    r = b["ranges"]
    nz = r[r[:, 1] > r[:, 0]]
    order = np.argsort(nz[:, 0])
    nz = nz[order]
    assert nz[0, 0] == 0 and nz[-1, 1] == b["P"]
    np.testing.assert_array_equal(nz[1:, 0], nz[:-1, 1])


def test_empty_and_all_culled():
    b = binning.bin_gaussians(np.zeros((4, 2), np.float32), np.zeros(4, np.int32), 32, 32)
    assert b["P"] == 0
    np.testing.assert_array_equal(b["point_list"], np.zeros(1, np.int32))
    assert (b["ranges"] == 0).all()
    # Gaussians far outside the image are clamped to an empty rectangle
    b = binning.bin_gaussians(np.array([[-500, -500], [900, 10]], np.float32),
                              np.array([3, 3], np.int32), 32, 32)
    assert b["P"] == 0


def test_hand_worked_rectangles():
    W = H = 64   # 4x4 tiles
    means = np.array([[16.0, 16.0],     # r=1: x in (15..17) -> lo 0, hi (16+1+15)/16=2 -> tiles 0..1
                      [0.5, 0.5],       # r=1: lo trunc(-0.5/16)=0, hi trunc(16.5/16)=1
                      [63.9, 63.9],     # r=2: lo trunc(61.9/16)=3, hi min(4, trunc(80.9/16)=5)=4
                      [32.0, 8.0]],     # r=0: culled
                     np.float32)
    radii = np.array([1, 1, 2, 0], np.int32)
    lo_x, lo_y, hi_x, hi_y = binning.get_rect(means, radii, W, H)
    np.testing.assert_array_equal(lo_x, [0, 0, 3, 2])
    np.testing.assert_array_equal(hi_x, [2, 1, 4, 2])
    np.testing.assert_array_equal(binning.tiles_touched(means, radii, W, H), [4, 1, 1, 0])
    b = binning.bin_gaussians(means, radii, W, H)
    np.testing.assert_array_equal(b["point_list"], [0, 1, 0, 0, 0, 2])
    assert tuple(b["ranges"][0]) == (0, 2) and tuple(b["ranges"][15]) == (5, 6)


def test_float32_rounding_of_rect_upper_bound():
    # pos + r crosses into the next binade, so ((pos + r) + 16) - 1 rounds in float32;
    # the oracle must follow the float32 sequence, not exact arithmetic.
    pos = np.float32(1000.00006)
    r = 40
    v = ((pos + np.float32(r)) + np.float32(16)) - np.float32(1)
    lo, lo_y, hi, hi_y = binning.get_rect(np.array([[pos, pos]], np.float32), np.array([r]), 2048, 2048)
    assert hi[0] == int(np.trunc(v / np.float32(16)))


def test_sub_bins_chunks_reproduce_the_whole_frame():
    """Rendering a frame a few tile rows at a time from sub_bins gives the whole render."""
    import torch
    from oracle import raster
    from oracle.step import render_binned_diff
    from test_oracle_raster import make_inputs
    W, H = 72, 56
    d = make_inputs(300, W, H, 3, 9)
    full, bins = raster.render_binned(d["means"], d["inv_cov"], d["colours"], d["phase"],
                                      d["opac"], d["curv"], d["area"], d["radii"], d["k"],
                                      d["dxy2"], W, H)
    ri = {k: torch.as_tensor(d[n]).double() for k, n in
          (("means", "means"), ("inv_cov", "inv_cov"), ("colours", "colours"),
           ("phase", "phase"), ("opac", "opac"), ("curv", "curv"), ("area", "area"))}
    ri["k"], ri["dxy2"] = torch.as_tensor(d["k"]).double(), d["dxy2"]
    gx, gy = binning.tile_grid(W, H)
    acc = torch.zeros_like(full)
    for ty0 in range(0, gy, 2):
        tiles = [ty * gx + tx for ty in range(ty0, min(gy, ty0 + 2)) for tx in range(gx)]
        acc += render_binned_diff(ri, W, H, binning.sub_bins(bins, tiles))
    assert float((acc - full).abs().max()) <= 1e-12
