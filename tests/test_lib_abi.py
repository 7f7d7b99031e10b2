"""CPU checks of the C ABI: the library loads and exports every symbol the header declares."""
import ctypes
import os
import re

import pytest

from conftest import PKG_DIR, ROOT

HEADER = os.path.join(ROOT, "include", "cvqpg.h")
LIB = os.path.join(PKG_DIR, "libcvqpg.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cvqpg_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def handle():
    if not os.path.exists(LIB):
        pytest.fail("libcvqpg.so is not built (run __graft_entry__.build())")
    import torch  # noqa: F401  -- same HIP runtime load order as the package
    return ctypes.CDLL(LIB)


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("cvqpg_last_error", "cvqpg_bin_count", "cvqpg_bin_sort", "cvqpg_render_forward",
              "cvqpg_render_backward", "cvqpg_qpf_render_forward"):
        assert s in syms


def test_library_exports_every_declared_symbol(handle):
    missing = [s for s in declared_symbols() if not hasattr(handle, s)]
    assert not missing, missing


def test_python_binding_covers_header(pkg):
    from cvqpg_amd import _lib
    assert set(declared_symbols()) == set(_lib.SIGNATURES)
    L = _lib.lib()
    assert L.cvqpg_abi_version() == 2


def test_workspace_size_and_argument_errors(pkg):
    from cvqpg_amd import _lib
    L = _lib.lib()
    small = L.cvqpg_raster_workspace_size(1000, 3, 64, 64, 0)
    big = L.cvqpg_raster_workspace_size(1000, 3, 64, 64, 100000)
    assert 0 < small < big
    assert big >= 100000 * 64  # one 64-byte partial record per key
    # invalid channel count is rejected on the host before any GPU work
    rc = L.cvqpg_render_forward(*([None] * 8), 10, 2, 1e-10, 64, 64, None, 0, 0, None, None,
                                None, None)
    assert rc == 1
    assert b"1 or 3 channels" in L.cvqpg_last_error()
    rc = L.cvqpg_bin_count(None, None, 10, 64, 64, None, 0, 0, None, None)
    assert rc == 1 and b"workspace" in L.cvqpg_last_error()
    rc = L.cvqpg_bin_count_live(None, None, 10, 64, 64, None, 0, 0, None, None)
    assert rc == 1 and b"workspace" in L.cvqpg_last_error()
    rc = L.cvqpg_bin_scan(10, None, 0, 0, None, None)
    assert rc == 1 and b"workspace" in L.cvqpg_last_error()
    nb = L.cvqpg_raster_workspace_size(10, 3, 64, 64, 100)
    rc = L.cvqpg_prologue_forward_live(None, 10, 3, 64, 64, None, 1e-10, 5e-7, 2, 0.9, 1, nb,
                                       100, None, None, None)
    assert rc == 1 and b"NULL" in L.cvqpg_last_error()
    # target window statistics: (D C) x (H - 10) x (W - 10) float2, checked before any launch
    assert L.cvqpg_loss_target_stats_size(3, 4, 40, 56) == 4 * 3 * 30 * 46 * 8
    assert L.cvqpg_loss_target_stats_size(3, 4, 10, 56) == 0
    rc = L.cvqpg_loss_target_stats(None, 3, 4, 40, 56, None, 0, None)
    assert rc == 1 and b"NULL" in L.cvqpg_last_error()
    rc = L.cvqpg_loss_target_stats(1, 3, 4, 40, 56, 1, 100, None)
    assert rc == 1 and b"too small" in L.cvqpg_last_error()
    # BLASM pass masks: 1 | 2 | 4, CVQPG_BLASM_TABLES (8) alone or with passes,
    # CVQPG_BLASM_TABLES_KEPT (16) only with something to run
    wl, dist = _lib.doubles([639e-9, 515e-9, 473e-9]), _lib.doubles([0.0, 2e-3])
    for bad in (0, 16, 32, -1):
        rc = L.cvqpg_blasm_forward_passes(bad, None, 3, 40, 56, 64, 128, 2, wl, dist, 8e-6,
                                          0.0, 0, None, 0, None, None, None)
        assert rc == 1 and b"mask" in L.cvqpg_last_error(), bad
        rc = L.cvqpg_blasm_backward_passes(bad, None, None, None, 3, 40, 56, 64, 128, 2, wl,
                                           dist, 8e-6, 0.0, 0, None, 0, None, None)
        assert rc == 1 and b"mask" in L.cvqpg_last_error(), bad
    # a tables-only call needs the workspace but no data pointers
    rc = L.cvqpg_blasm_forward_passes(8, None, 3, 40, 56, 64, 128, 2, wl, dist, 8e-6, 0.0, 0,
                                      None, 0, None, None, None)
    assert rc == 1 and b"workspace" in L.cvqpg_last_error()
    rc = L.cvqpg_blasm_forward_passes(1 | 16, None, 3, 40, 56, 64, 128, 2, wl, dist, 8e-6, 0.0,
                                      0, None, 0, None, None, None)
    assert rc == 1 and b"NULL" in L.cvqpg_last_error()


def test_batch_table_layout_and_errors(pkg):
    """cvqpg_raster_batch_table is host-only: scenes in queue order, tile offsets, and the
    argument checks (no GPU needed, fake device addresses)."""
    import numpy as np
    from cvqpg_amd import _lib
    L = _lib.lib()
    sizes = [(1920, 1080, 5000), (512, 512, 100), (100, 36, 7)]
    scenes = (_lib.RasterScene * 3)()
    for i, (W, H, n) in enumerate(sizes):
        s = scenes[i]
        s.n, s.width, s.height, s.channels = n, W, H, 3
        s.capacity = 4 * n
        s.ws = 0x10_0000_0000 * (i + 1)
        s.ws_bytes = L.cvqpg_raster_workspace_size(n, 3, W, H, s.capacity)
        s.point_list, s.ranges, s.field = s.ws + 0x1000_0000, s.ws + 0x2000_0000, s.ws + 0x3000_0000
        s.grad_field, s.k = s.ws + 0x4000_0000, s.ws + 0x5000_0000
    nb = L.cvqpg_raster_batch_table_size(3)
    assert nb == 64 + 3 * 96
    buf = (ctypes.c_uint8 * nb)()
    _lib.check(L.cvqpg_raster_batch_table(scenes, 3, buf, nb))
    raw = bytes(buf)
    hdr = np.frombuffer(raw[:64], dtype=np.int32)
    tiles = [((W + 15) // 16) * ((H + 15) // 16) for W, H, _ in sizes]
    assert list(hdr[:4]) == [0, 0, 0, 0]
    assert hdr[4] == 3 and hdr[5] == sum(tiles)
    begin = 0
    for i, (W, H, n) in enumerate(sizes):
        rec = raw[64 + 96 * i: 64 + 96 * (i + 1)]
        ptrs = np.frombuffer(rec[:80], dtype=np.uint64)
        ints = np.frombuffer(rec[80:], dtype=np.int32)
        s = scenes[i]
        # ranges, point_list, packed (= workspace base), rect, misc, field, grad, k, partials,
        # kept-tile masks of the live binning
        assert ptrs[0] == s.ranges and ptrs[1] == s.point_list and ptrs[2] == s.ws
        assert s.ws < ptrs[3] < ptrs[9] < ptrs[4] < ptrs[8] < s.ws + s.ws_bytes
        assert ptrs[5] == s.field and ptrs[6] == s.grad_field and ptrs[7] == s.k
        assert list(ints) == [W, H, (W + 15) // 16, begin]
        begin += tiles[i]
    # errors: mixed channels, workspace too small, table too small, scene count
    scenes[2].channels = 1
    assert L.cvqpg_raster_batch_table(scenes, 3, buf, nb) == 1
    assert b"channels" in L.cvqpg_last_error()
    scenes[2].channels = 3
    scenes[1].ws_bytes = 100
    assert L.cvqpg_raster_batch_table(scenes, 3, buf, nb) == 1
    assert b"workspace too small" in L.cvqpg_last_error()
    scenes[1].ws_bytes = L.cvqpg_raster_workspace_size(100, 3, 512, 512, 400)
    assert L.cvqpg_raster_batch_table(scenes, 3, buf, nb - 1) == 1
    assert L.cvqpg_raster_batch_table(scenes, 0, buf, nb) == 1
    assert L.cvqpg_raster_batch_table(scenes, 65, buf, nb) == 1
    assert L.cvqpg_render_forward_batch(None, 3, 256, 1, None) == 1
    assert L.cvqpg_render_backward_batch(buf, 2, 256, 1, None) == 1
    assert L.cvqpg_render_forward_batch(buf, 3, 256, 0, None) == 1


def test_overflow_flags_point_inside_the_workspace(pkg):
    """cvqpg_raster_overflow_flags is host arithmetic only (no device access), so it can be
    checked with a made-up workspace address."""
    import ctypes
    from cvqpg_amd import _lib
    L = _lib.lib()
    n, cap = 5000, 70000
    nb = L.cvqpg_raster_workspace_size(n, 3, 640, 480, cap)
    base = 1 << 40
    out = ctypes.c_void_p()
    assert L.cvqpg_raster_overflow_flags(ctypes.c_void_p(base), nb, n, cap, ctypes.byref(out)) == 0
    off = out.value - base
    assert 0 < off and off + 8 <= nb and off % 4 == 0
    # the flags precede every capacity-dependent region: a larger capacity keeps them in place
    nb2 = L.cvqpg_raster_workspace_size(n, 3, 640, 480, 2 * cap)
    out2 = ctypes.c_void_p()
    assert L.cvqpg_raster_overflow_flags(ctypes.c_void_p(base), nb2, n, 2 * cap,
                                         ctypes.byref(out2)) == 0
    assert out2.value == out.value
    assert L.cvqpg_raster_overflow_flags(ctypes.c_void_p(base), nb - 1, n, cap,
                                         ctypes.byref(out)) == 1
