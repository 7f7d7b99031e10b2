"""Host-side helpers of the package that run without a GPU.

``qpf_render_reference`` (the dense testing helper of cuda/rasterizer.py:74-90) against
oracle.raster.render_dense, in FP32 and FP64, with row bands small enough that several
are evaluated, and its autograd gradient against the oracle's closed-form adjoint.
"""
import numpy as np
import pytest
import torch

from oracle import binning, raster
from test_oracle_raster import make_inputs

NAMES = ("means", "inv_cov", "colours", "phase", "opac", "curv", "area")


@pytest.mark.parametrize("n,W,H,C,dt,bar", [(50, 30, 20, 3, torch.float32, 1e-6),
                                            (40, 17, 23, 1, torch.float64, 1e-12)])
def test_dense_reference_matches_oracle(pkg, n, W, H, C, dt, bar):
    d = make_inputs(n, W, H, C, 3, curv_scale=0.05)
    t = {k: torch.as_tensor(d[k]).to(dt) for k in NAMES + ("k",)}
    got = pkg.qpf_render_reference(*(t[k] for k in NAMES), t["k"], d["dxy2"], W, H,
                                   pixels_per_chunk=2000)
    ref = raster.render_dense(*(d[k] for k in NAMES), d["k"], d["dxy2"], W, H)
    assert got.shape == (C, H, W)
    e = float((got.to(torch.complex128) - ref).abs().norm() / ref.abs().norm())
    assert e <= bar, e


def test_dense_reference_gradient_matches_closed_form_adjoint(pkg):
    n, W, H, C = 30, 20, 18, 3
    d = make_inputs(n, W, H, C, 5, curv_scale=0.05, smin=0.5, smax=2.0)
    leaves = {k: torch.as_tensor(d[k]).double().requires_grad_() for k in NAMES}
    out = pkg.qpf_render_reference(*(leaves[k] for k in NAMES), torch.as_tensor(d["k"]).double(),
                                   d["dxy2"], W, H, pixels_per_chunk=700)
    rng = np.random.default_rng(6)
    G = torch.complex(torch.as_tensor(rng.normal(size=(C, H, W))),
                      torch.as_tensor(rng.normal(size=(C, H, W))))
    (G.real * out.real + G.imag * out.imag).sum().backward()
    # every pair inside the tile rectangles of radii this large: the binned adjoint is dense
    radii = np.full(n, 64, np.int32)
    bins = binning.bin_gaussians(d["means"], radii, W, H)
    ref = raster.render_backward(G, *(d[k] for k in NAMES), d["k"], d["dxy2"], W, H, bins)
    for k, r in zip(NAMES, ref):
        e = float((leaves[k].grad - r).norm() / r.norm())
        assert e <= 1e-6, (k, e)
