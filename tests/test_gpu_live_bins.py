"""Live binning of the fused step (cvqpg_prologue_forward_live + cvqpg_bin_scan, and
cvqpg_bin_count_live) against the reference's full tile lists.

The fused step keys a tile of a Gaussian's 5-sigma rectangle only if the Gaussian's exponent
reaches the raster kernels' skip threshold (2^-24 of a unit term) less a margin somewhere on
the tile; both tile kernels skip every other key.  Checks:
  * the prologue's fused live count gives the lists of cvqpg_bin_count_live bit for bit;
  * each tile's live list is an ordered subsequence of the reference's full list
    (oracle.binning, the reference's 64-bit key sort), and every key it leaves out belongs to
    a Gaussian whose FP64 exponent stays below 2^-24 on all 256 pixels of that tile;
  * the fused step on live lists gives the loss, the raw gradients and (after the Adan
    update) the parameters of the same step on full lists bit for bit, over three steps.
"""
import numpy as np
import pytest
import torch

from oracle import binning
from oracle import model as omodel
from oracle import step as ostep

pytestmark = pytest.mark.gpu
DEV = "cuda"
LOG2E = 1.4426950408889634
SPEC_SMALL = dict(W=64, H=48, N=300, wavelengths=[639e-9, 515e-9, 473e-9], pitch=8e-6,
                  distances=[0.0, 3e-3, 6e-3], kappa=0.05)
# 32 x 32 = 1024 tiles: 10 key bits, two radix passes
SPEC_512 = dict(W=512, H=512, N=20000, wavelengths=[639e-9, 515e-9, 473e-9], pitch=8e-6,
                distances=[0.0, 2e-3], kappa=0.05)
# partial tiles at the right (8 px) and bottom (6 px) edges
SPEC_EDGE = dict(W=200, H=150, N=2000, wavelengths=[639e-9, 515e-9, 473e-9], pitch=8e-6,
                 distances=[0.0, 3e-3], kappa=0.05)


def trainer(spec, full=False, seed=0):
    from cvqpg_amd import _lib, fused, train
    wl = train.Workload("test", device=DEV, seed=seed, spec=spec)
    tr = fused.FusedTrainer(wl, num_itrs=100, warmup_iters=0, inject_noise=False)
    if full:
        def full_count():
            ws, nb, cap = tr._ws()
            _lib.check(_lib.lib().cvqpg_bin_count(
                _lib.ptr(tr.means), _lib.ptr(tr.radii), tr.N, tr.W, tr.H, ws, nb, cap,
                _lib.ptr(tr.num_pairs), tr._s()))
        tr._count = full_count
        tr.grow()  # key buffers for the full lists
    return wl, tr


@pytest.mark.parametrize("spec", [SPEC_SMALL, SPEC_EDGE, SPEC_512], ids=["64x48", "200x150", "512x512"])
def test_live_lists_are_the_kernels_lists(pkg, spec):
    wl, tr = trainer(spec)
    st = dict(tr._stages())
    for name in ("prologue", "bin_count", "bin_sort"):
        st[name]()
    torch.cuda.synchronize()
    W, H = wl.img_size
    P = int(tr.num_pairs.item())
    assert 0 < P <= tr.capacity
    pl = tr.point_list[:P].cpu().numpy()
    rg = tr.ranges.cpu().numpy()
    # the prologue's fused live count equals cvqpg_bin_count_live's, bit for bit
    from cvqpg_amd import _lib
    ws, nb, cap = tr._ws()
    _lib.check(_lib.lib().cvqpg_bin_count_live(
        _lib.ptr(tr.means), _lib.ptr(tr.radii), tr.N, tr.W, tr.H, ws, nb, cap,
        _lib.ptr(tr.num_pairs), tr._s()))
    st["bin_sort"]()
    torch.cuda.synchronize()
    assert int(tr.num_pairs.item()) == P
    assert np.array_equal(tr.point_list[:P].cpu().numpy(), pl)
    assert np.array_equal(tr.ranges.cpu().numpy(), rg)
    means = tr.means.cpu().numpy()
    radii = tr.radii.cpu().numpy()
    full = binning.bin_gaussians(means, radii, W, H)
    assert P < full["P"]
    print(f"{W}x{H}: {P} live keys of {full['P']} ({P / full['P']:.3f})", flush=True)
    raw = {n: getattr(wl.gaussians, n).detach().cpu().double() for n in ostep.adan.PARAM_NAMES}
    ap = wl.args_prop
    ri = omodel.raster_inputs(raw, wl.img_size, ap.wavelengths, ap.pixel_pitch)
    m64 = ri["means"].numpy()
    ic = ri["inv_cov"].numpy()
    gx = (W + 15) // 16
    rem_g, rem_t = [], []
    for t in range(rg.shape[0]):
        live = pl[rg[t, 0]:rg[t, 1]]
        ref = full["point_list"][full["ranges"][t, 0]:full["ranges"][t, 1]]
        assert np.all(np.diff(live) > 0), t
        assert np.isin(live, ref).all(), t
        out = np.setdiff1d(ref, live)
        rem_g.append(out)
        rem_t.append(np.full(out.size, t))
    rem_g, rem_t = np.concatenate(rem_g), np.concatenate(rem_t)
    assert rem_g.size == full["P"] - P
    # every left-out key: FP64 exponent below 2^-24 on every pixel of its tile
    yy, xx = np.mgrid[0:16, 0:16]
    worst = -np.inf
    for s in range(0, rem_g.size, 4096):
        g, t = rem_g[s:s + 4096], rem_t[s:s + 4096]
        px = (t % gx)[:, None] * 16 + xx.ravel()[None, :]
        py = (t // gx)[:, None] * 16 + yy.ravel()[None, :]
        inside = (px < W) & (py < H)
        dx = m64[g, 0][:, None] - px
        dy = m64[g, 1][:, None] - py
        mahal = ic[g, 0][:, None] * dx * dx + 2 * ic[g, 1][:, None] * dx * dy + \
            ic[g, 2][:, None] * dy * dy
        q2 = np.where(inside, -0.5 * LOG2E * mahal, -np.inf)
        worst = max(worst, float(q2.max()))
    print(f"largest log2 G of a left-out key: {worst:.3f}", flush=True)
    assert worst < -24.0


@pytest.mark.parametrize("spec", [SPEC_EDGE, SPEC_512], ids=["200x150", "512x512"])
def test_live_step_equals_full_step(pkg, spec):
    wl_a, a = trainer(spec)
    wl_b, b = trainer(spec, full=True)
    for pa, pb in zip(a.params, b.params):
        assert torch.equal(pa, pb)
    for it in range(3):
        a.step()
        b.step()
        torch.cuda.synchronize()
        assert torch.equal(a.loss, b.loss), it
        assert torch.equal(a.grads, b.grads), it
        for pa, pb in zip(a.params, b.params):
            assert torch.equal(pa, pb), it
    assert a.check_capacity() and b.check_capacity()
    assert a.last_pairs < b.last_pairs
