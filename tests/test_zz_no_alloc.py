"""The no-allocation loss entry point (static device scratch) against the workspace entry
point it shares its kernels with: the same loss and gradient bit for bit, on two calls in
a row, the second with a smaller frame, both with and without precomputed target
statistics and target noise; sizes beyond the static scratch are rejected on the host."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _inputs(C, D, H, W, seed):
    rng = np.random.default_rng(seed)
    f = lambda *s: torch.as_tensor(rng.uniform(-0.1, 1.1, s), dtype=torch.float32).to(DEV)
    inten = f(D, C, H, W)
    target = f(D, C, H, W).clamp(0, 1)
    mask = torch.as_tensor(rng.uniform(0, 1, (D, C, H, W)) > 0.5, dtype=torch.float32).to(DEV)
    return inten, target, mask


def _run(L, _lib, entry, inten, target, mask, stats, ns, step):
    D, C, H, W = inten.shape
    grad = torch.empty_like(inten)
    loss = torch.zeros(1, device=DEV)
    nsp = None if ns is None else _lib.floats(ns)
    sp = _lib.ptr(stats)
    s = _lib.stream_handle()
    if entry == "ws":
        nb = int(L.cvqpg_loss_workspace_size(C, D, H, W))
        ws = torch.empty(nb, dtype=torch.uint8, device=DEV)
        _lib.check(L.cvqpg_loss_forward_backward_stats(
            _lib.ptr(inten), _lib.ptr(target), sp, _lib.ptr(mask), None, nsp, 123,
            _lib.ptr(step), C, D, H, W, _lib.ptr(ws), nb, _lib.ptr(grad), _lib.ptr(loss), s))
    else:
        _lib.check(L.cvqpg_loss_forward_backward_noalloc(
            _lib.ptr(inten), _lib.ptr(target), sp, _lib.ptr(mask), None, nsp, 123,
            _lib.ptr(step), C, D, H, W, _lib.ptr(grad), _lib.ptr(loss), s))
    torch.cuda.synchronize()
    return loss, grad


@pytest.mark.parametrize("with_stats", [False, True])
def test_noalloc_loss_equals_workspace_loss(pkg, with_stats):
    from cvqpg_amd import _lib
    L = _lib.lib()
    step = torch.tensor([7], dtype=torch.int32, device=DEV)
    # 1080p-shaped first call, then a smaller frame: the second call must not see the
    # first call's partials
    for (C, D, H, W), seed in (((3, 4, 1080, 1920), 0), ((1, 3, 256, 256), 1)):
        inten, target, mask = _inputs(C, D, H, W, seed)
        stats = None
        if with_stats:
            nts = int(L.cvqpg_loss_target_stats_size(C, D, H, W))
            stats = torch.empty(nts // 4, device=DEV)
            _lib.check(L.cvqpg_loss_target_stats(_lib.ptr(target), C, D, H, W, _lib.ptr(stats),
                                                 nts, _lib.stream_handle()))
        ns = [1e-3 * (d + 1) for d in range(D)]
        lw, gw = _run(L, _lib, "ws", inten, target, mask, stats, ns, step)
        ln, gn = _run(L, _lib, "noalloc", inten, target, mask, stats, ns, step)
        print(f"C={C} D={D} {W}x{H} stats={with_stats}: loss {float(lw):.7e} / {float(ln):.7e}",
              flush=True)
        assert torch.equal(lw, ln)
        assert torch.equal(gw, gn)


def test_noalloc_rejects_sizes_beyond_its_scratch(pkg):
    from cvqpg_amd import _lib
    L = _lib.lib()
    assert int(L.cvqpg_loss_noalloc_max_bytes()) == 256 * 1024
    # 16 planes x 3 channels of 8192 x 8192 need far more partial sums than the scratch holds;
    # the check happens on the host before any pointer is used
    rc = L.cvqpg_loss_forward_backward_noalloc(None, None, None, None, None, None, 0, None,
                                               3, 16, 8192, 8192, None, None, None)
    assert rc == 1
    assert b"no-allocation" in L.cvqpg_last_error()
