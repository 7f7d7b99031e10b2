"""GPU parity of the warm-start network kernels (bf16 MFMA, FP32 accumulation) against the
CPU restatement in oracle/nn.py.

Bars: each kernel on bf16-exact inputs within 1e-5 relative (only the FP32 summation
order differs); the whole network, whose activations are stored in bf16, within 2e-2
relative L2 of the oracle's forward and 5e-2 of its FP64 gradients per layer (bf16
rounding of the pre-activation gradients is the GPU's only other approximation).
"""
import math

import pytest
import torch

from oracle import nn as onn

pytestmark = pytest.mark.gpu
DEV = "cuda"


def lib():
    from cvqpg_amd import _lib
    return _lib


def rel(a, b):
    a, b = a.double().cpu(), b.double().cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-30))


def rnd_bf16(shape, seed, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    return (torch.randn(*shape, generator=g) * scale).to(torch.bfloat16)


@pytest.mark.parametrize("k,B,H,W,cin,cout", [(3, 2, 5, 7, 32, 48), (1, 1, 3, 4, 1024, 32),
                                              (3, 3, 9, 6, 64, 112), (3, 1, 1, 1, 256, 256),
                                              (3, 2, 37, 45, 32, 32)])  # 3 pixel chunks
def test_conv_forward_wgrad_dgrad(pkg, k, B, H, W, cin, cout):
    _lib = lib()
    L = _lib.lib()
    x = rnd_bf16((B, H, W, cin), 1)
    w = rnd_bf16((cout, k * k * cin), 2, 1.0 / math.sqrt(cin * k * k))
    b = torch.randn(cout, generator=torch.Generator().manual_seed(3))
    xd, wd, bd = x.to(DEV), w.to(DEV), b.to(DEV)
    yf = torch.empty(B, H, W, cout, device=DEV)
    yb = torch.empty(B, H, W, cout, dtype=torch.bfloat16, device=DEV)
    s = _lib.stream_handle()
    _lib.check(L.cvqpg_nn_conv(_lib.ptr(xd), B, H, W, cin, _lib.ptr(wd), _lib.ptr(bd), cout, k, 1,
                               None, _lib.ptr(yb), None, _lib.ptr(yf), s))
    ref = torch.relu(onn.conv(x.double(), w.double(), b.double(), k))
    e = rel(yf, ref)
    print(f"conv k={k} {B}x{H}x{W} {cin}->{cout}: forward rel {e:.1e}", flush=True)
    assert e <= 1e-5
    mism = (yb.cpu().float() != onn.bf16(ref).float()).float().mean().item()
    assert mism <= 0.01
    # weight gradient and data gradient of y = conv(x, w) for a bf16 upstream gradient
    dy = rnd_bf16((B, H, W, cout), 4)
    dyd = dy.to(DEV)
    dw = torch.empty(cout, k * k * cin, device=DEV)
    db = torch.empty(cout, device=DEV)
    wsz = int(L.cvqpg_nn_wgrad_workspace_size(B, H, W, cin, cout, k))
    wws = torch.empty(wsz, dtype=torch.uint8, device=DEV)
    _lib.check(L.cvqpg_nn_conv_wgrad(_lib.ptr(xd), B, H, W, cin, _lib.ptr(dyd), cout, k,
                                     _lib.ptr(dw), _lib.ptr(db), _lib.ptr(wws), wsz, s))
    x64 = x.double().requires_grad_()
    w64 = w.double().requires_grad_()
    (onn.conv(x64, w64, None, k) * dy.double()).sum().backward()
    ew = rel(dw, w64.grad)
    wf = torch.empty(cin * k * k * cout, dtype=torch.bfloat16, device=DEV)
    _lib.check(L.cvqpg_nn_flip_filter(_lib.ptr(wd), cout, cin, k, _lib.ptr(wf), s))
    dx = torch.empty(B, H, W, cin, device=DEV)
    _lib.check(L.cvqpg_nn_conv(_lib.ptr(dyd), B, H, W, cout, _lib.ptr(wf), None, cin, k, 0, None,
                               None, None, _lib.ptr(dx), s))
    ex = rel(dx, x64.grad)
    eb = rel(db, dy.double().sum((0, 1, 2)))
    print(f"  wgrad rel {ew:.1e}, dgrad rel {ex:.1e}, bias rel {eb:.1e}", flush=True)
    assert ew <= 1e-5 and ex <= 1e-5 and eb <= 1e-5
    # accumulate flag and the skip-sum output
    _lib.check(L.cvqpg_nn_conv(_lib.ptr(dyd), B, H, W, cout, _lib.ptr(wf), None, cin, k, 2, None,
                               None, None, _lib.ptr(dx), s))
    assert rel(dx, 2 * x64.grad) <= 1e-5
    add = rnd_bf16((B, H, W, cout), 5).to(DEV)
    y2 = torch.empty_like(yb)
    _lib.check(L.cvqpg_nn_conv(_lib.ptr(xd), B, H, W, cin, _lib.ptr(wd), _lib.ptr(bd), cout, k, 1,
                               _lib.ptr(add), _lib.ptr(yb), _lib.ptr(y2), None, s))
    exp = onn.bf16(yb.cpu().float() + add.cpu().float())
    assert torch.equal(y2.cpu().float(), exp)


def test_pool_upsample_tiles(pkg):
    _lib = lib()
    L = _lib.lib()
    s = _lib.stream_handle()
    B, H, W, C = 2, 5, 7, 32
    x = rnd_bf16((B, H, W, C), 6)
    y = torch.empty(B, 3, 4, C, dtype=torch.bfloat16, device=DEV)
    _lib.check(L.cvqpg_nn_pool2(_lib.ptr(x.to(DEV)), B, H, W, C, _lib.ptr(y), s))
    ref = onn.bf16(onn.pool2(x.double()))
    assert (y.cpu().double() - ref).abs().max() <= 1e-2 * ref.abs().max()
    assert (y.cpu().double() != ref).float().mean() <= 0.01
    g = torch.randn(B, 3, 4, C)
    dx = torch.zeros(B, H, W, C, device=DEV)
    _lib.check(L.cvqpg_nn_pool2_backward(_lib.ptr(g.to(DEV)), B, H, W, C, _lib.ptr(dx), 1, s))
    x64 = x.double().requires_grad_()
    (onn.pool2(x64) * g.double()).sum().backward()
    assert rel(dx, x64.grad) <= 1e-6
    up = torch.empty(B, H, W, C, dtype=torch.bfloat16, device=DEV)
    _lib.check(L.cvqpg_nn_upsample2(_lib.ptr(y), B, 3, 4, C, H, W, _lib.ptr(up), s))
    assert torch.equal(up.cpu(), onn.up2(y.cpu(), H, W))
    gu = torch.randn(B, H, W, C)
    dsrc = torch.empty(B, 3, 4, C, device=DEV)
    _lib.check(L.cvqpg_nn_upsample2_backward(_lib.ptr(gu.to(DEV)), B, 3, 4, C, H, W,
                                             _lib.ptr(dsrc), 0, s))
    s64 = y.cpu().double().requires_grad_()
    (onn.up2(s64, H, W) * gu.double()).sum().backward()
    assert rel(dsrc, s64.grad) <= 1e-6
    img = torch.rand(2, 4, 40, 35)
    t = torch.empty(2, 3, 3, 1024, dtype=torch.bfloat16, device=DEV)
    _lib.check(L.cvqpg_nn_tiles_from_image(_lib.ptr(img.to(DEV)), 2, 4, 40, 35, _lib.ptr(t), s))
    assert torch.equal(t.cpu().float(), onn.bf16(onn.tiles_from_image(img)))


def test_adam_and_wmse(pkg):
    _lib = lib()
    L = _lib.lib()
    s = _lib.stream_handle()
    n = 1000
    g = torch.Generator().manual_seed(7)
    w = torch.randn(n, generator=g)
    m, v = torch.zeros(n), torch.zeros(n)
    wd, md, vd = w.to(DEV), m.to(DEV), v.to(DEV)
    wb = torch.empty(n, dtype=torch.bfloat16, device=DEV)
    for step in (1, 2, 3):
        gr = torch.randn(n, generator=g)
        _lib.check(L.cvqpg_nn_adam(_lib.ptr(wd), _lib.ptr(wb), _lib.ptr(gr.to(DEV)), _lib.ptr(md),
                                   _lib.ptr(vd), n, 1e-2, 0.9, 0.999, 1e-8, step, s))
        m = 0.9 * m + 0.1 * gr
        v = 0.999 * v + 0.001 * gr * gr
        w = w - 1e-2 * (m / (1 - 0.9 ** step)) / (torch.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
    assert rel(wd, w) <= 1e-6
    assert torch.equal(wb.cpu(), wd.cpu().to(torch.bfloat16))
    out, tgt, wt = torch.randn(3, 5, 112), torch.randn(3, 5, 112), torch.rand(3, 5, 112)
    ws = torch.empty(int(L.cvqpg_nn_wmse_workspace_size(out.numel())), dtype=torch.uint8, device=DEV)
    dout = torch.empty(3, 5, 112, device=DEV)
    loss = torch.empty(1, device=DEV)
    od, td, wtd = out.to(DEV), tgt.to(DEV), wt.to(DEV)   # kept alive across the call
    _lib.check(L.cvqpg_nn_wmse(_lib.ptr(od), _lib.ptr(td), _lib.ptr(wtd), out.numel(), 0.25,
                               _lib.ptr(dout), _lib.ptr(ws), ws.numel(), _lib.ptr(loss), s))
    assert abs(float(loss) - float((0.25 * wt * (out - tgt) ** 2).sum())) <= 1e-4
    assert rel(dout, 0.5 * wt * (out - tgt)) <= 1e-6


def test_network_forward_and_gradients_match_oracle(pkg):
    from cvqpg_amd.warmstart import OUT_DIM, WarmStartNet
    net = WarmStartNet(DEV, seed=11)
    params = net.state_dict()
    # nonzero biases so that every bias path is exercised
    g = torch.Generator().manual_seed(12)
    params = {n: (w, 0.05 * torch.randn(b.shape, generator=g)) for n, (w, b) in params.items()}
    net.load_state_dict(params)
    img = torch.rand(2, 4, 72, 120, generator=g)
    out = net.forward(img.to(DEV))
    torch.cuda.synchronize()
    target = torch.randn(2, 5, 8, OUT_DIM, generator=g)
    weight = torch.rand(2, 5, 8, OUT_DIM, generator=g)
    scale = 1.0 / target.numel()
    Lr, ref_out, ref_grads = onn.loss_and_grads(img, params, target, weight, scale)
    eo = rel(out, ref_out)
    print(f"network forward rel L2 {eo:.2e}", flush=True)
    assert eo <= 2e-2
    A = net.activations(2, 5, 8)
    net.loss_backward(A, target.to(DEV), weight.to(DEV), scale)
    net.backward(A)
    torch.cuda.synchronize()
    assert abs(float(net.loss) - Lr) <= 2e-2 * abs(Lr)
    worst = 0.0
    for name, *_ in net.layers:
        gw, gb = net.view(net.grad, name)
        ew, eb = rel(gw, ref_grads[name][0]), rel(gb, ref_grads[name][1])
        worst = max(worst, ew, eb)
        print(f"  {name}: dW rel {ew:.1e}, db rel {eb:.1e}", flush=True)
        assert ew <= 5e-2 and eb <= 5e-2, name
    print(f"worst gradient rel L2 {worst:.2e}", flush=True)


def test_training_reduces_loss(pkg):
    """300 Adam steps on a fixed minibatch of random targets bring the loss below 30%."""
    from cvqpg_amd.warmstart import OUT_DIM, WarmStartNet
    net = WarmStartNet(DEV, seed=1)
    g = torch.Generator().manual_seed(2)
    img = torch.rand(4, 4, 64, 64, generator=g).to(DEV)
    target = (0.5 * torch.randn(4, 4, 4, OUT_DIM, generator=g)).to(DEV)
    weight = torch.ones_like(target)
    first = float(net.train_step(img, target, weight, 1e-3))
    for _ in range(300):
        loss = net.train_step(img, target, weight, 1e-3)
    last = float(loss)
    print(f"training loss {first:.4f} -> {last:.4f}", flush=True)
    assert last < 0.3 * first


def test_encode_decode_round_trip(pkg):
    """decode(encode(fitted scene)) restores the strongest 8 Gaussians of each tile."""
    from cvqpg_amd import warmstart as ws
    from cvqpg_amd.model import PARAM_NAMES
    spec = ws.scene_spec(64, 48)
    wl = ws.make_workload(spec, 3, DEV)
    params = [getattr(wl.gaussians, n).detach() for n in PARAM_NAMES]
    t, w = ws.encode_targets(params, 64, 48)
    assert t.shape == (3, 4, ws.OUT_DIM)
    dec, present = ws.decode(t, 64, 48)
    a = ws.activated(params, 64, 48)
    b = ws.activated(dec, 64, 48)
    # every decoded Gaussian marked present matches one of the originals
    keep = present
    assert bool((b["opac"][~present] < 0.02).all())
    d = (b["mx"][keep, None] - a["mx"][None]) ** 2 + (b["my"][keep, None] - a["my"][None]) ** 2
    j = d.argmin(1)
    assert float(d.min(1).values.max()) < 1e-6
    assert torch.allclose(b["colours"][keep], a["colours"][j].abs(), atol=1e-5)
    assert torch.allclose(b["opac"][keep], a["opac"][j], atol=1e-5)
    inside = a["curv"][j].abs() < 0.99            # decode clamps |tanh| to 0.99
    assert torch.allclose(b["curv"][keep][inside], a["curv"][j][inside], atol=1e-5)
    tx = (a["mx"] / 16).floor().long().clamp(0, 3)
    ty = (a["my"] / 16).floor().long().clamp(0, 2)
    counts = torch.bincount(ty * 4 + tx, minlength=12)
    assert int(keep.sum()) == int(counts.clamp(max=8).sum())
