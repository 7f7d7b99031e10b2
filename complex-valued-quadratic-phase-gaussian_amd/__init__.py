"""cvqpg_amd -- MI355X-native CVQPG hologram hot path.

Hand-written gfx950 HIP kernels (libcvqpg.so, C ABI in include/cvqpg.h) behind
the reference's Python interfaces:

* ``rasterizer.qpf_render`` / ``QPFRender``   -- cuda/rasterizer.py drop-in
"""
from . import _lib  # noqa: F401
from .rasterizer import QPFRender, qpf_render, qpf_render_reference  # noqa: F401

__all__ = ["QPFRender", "qpf_render", "qpf_render_reference"]
