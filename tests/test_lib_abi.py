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
    assert L.cvqpg_abi_version() == 1


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
