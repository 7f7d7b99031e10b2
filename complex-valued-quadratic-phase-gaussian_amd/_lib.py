"""ctypes binding of libcvqpg.so (the C ABI declared in include/cvqpg.h).

torch is imported first on purpose: it loads its bundled HIP runtime
(libamdhip64.so.7), so the library's own dependency on that SONAME resolves to
the same runtime instance and both share devices, streams and allocations.

There is no fallback: if the library is missing or fails to load, every entry
point raises.
"""
import ctypes
import os

import torch  # noqa: F401  (must precede the library load)

HERE = os.path.dirname(os.path.abspath(__file__))
# CVQPG_LIB may point at another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("CVQPG_LIB", os.path.join(HERE, "libcvqpg.so"))

_c_int, _c_i64, _c_f32, _c_size = ctypes.c_int, ctypes.c_int64, ctypes.c_float, ctypes.c_size_t
_c_f64 = ctypes.c_double
_p = ctypes.c_void_p

# name -> (restype, argtypes); must match include/cvqpg.h
SIGNATURES = {
    "cvqpg_last_error": (ctypes.c_char_p, []),
    "cvqpg_abi_version": (_c_int, []),
    "cvqpg_raster_workspace_size": (_c_size, [_c_int, _c_int, _c_int, _c_int, _c_i64]),
    "cvqpg_bin_count": (_c_int, [_p, _p, _c_int, _c_int, _c_int, _p, _c_size, _c_i64, _p, _p]),
    "cvqpg_bin_count_live": (_c_int, [_p, _p, _c_int, _c_int, _c_int, _p, _c_size, _c_i64, _p, _p]),
    "cvqpg_bin_sort": (_c_int, [_p, _p, _c_int, _c_int, _c_int, _p, _c_size, _c_i64, _p, _p, _p]),
    "cvqpg_render_forward": (_c_int, [_p] * 8 + [_c_int, _c_int, _c_f32, _c_int, _c_int, _p,
                                                 _c_size, _c_i64, _p, _p, _p, _p]),
    "cvqpg_render_backward": (_c_int, [_p] * 9 + [_c_int, _c_int, _c_f32, _c_int, _c_int, _p,
                                                  _c_size, _c_i64, _p, _p] + [_p] * 7 + [_p]),
    "cvqpg_qpf_render_forward": (_c_int, [_p] * 9 + [_c_int, _c_int, _c_f32, _c_int, _c_int, _p,
                                                     _c_size, _c_i64, _p, _p, _p, _p, _p]),
    "cvqpg_raster_pack": (_c_int, [_p] * 8 + [_c_int, _c_int, _c_f32, _c_int, _c_int, _p, _c_size,
                                              _c_i64, _p]),
    "cvqpg_render_forward_packed": (_c_int, [_c_int, _c_int, _c_int, _c_int, _p, _c_size, _c_i64,
                                             _p, _p, _p, _p]),
    "cvqpg_render_backward_tiles": (_c_int, [_p, _p, _c_int, _c_int, _c_int, _c_int, _p, _c_size,
                                             _c_i64, _p, _p, _p]),
    "cvqpg_raster_batch_table_size": (_c_size, [_c_int]),
    "cvqpg_raster_batch_table": (_c_int, [_p, _c_int, _p, _c_size]),
    "cvqpg_render_forward_batch": (_c_int, [_p, _c_int, _c_int, _c_int, _p]),
    "cvqpg_render_backward_batch": (_c_int, [_p, _c_int, _c_int, _c_int, _p]),
    "cvqpg_render_backward_reduce": (_c_int, [_p, _p, _p, _c_int, _c_int, _c_f32, _c_int, _c_int,
                                              _p, _c_size, _c_i64] + [_p] * 7 + [_p]),
    "cvqpg_blasm_workspace_size": (_c_size, [_c_int] * 7),
    "cvqpg_blasm_forward": (_c_int, [_p] + [_c_int] * 6 + [_p, _p, _c_f64, _c_f64, _c_int, _p,
                                                          _c_size, _p, _p, _p]),
    "cvqpg_blasm_backward": (_c_int, [_p, _p, _p] + [_c_int] * 6 + [_p, _p, _c_f64, _c_f64, _c_int,
                                                                    _p, _c_size, _p, _p]),
    "cvqpg_blasm_forward_passes": (_c_int, [_c_int, _p] + [_c_int] * 6 + [_p, _p, _c_f64, _c_f64,
                                                                         _c_int, _p, _c_size, _p,
                                                                         _p, _p]),
    "cvqpg_blasm_backward_passes": (_c_int, [_c_int, _p, _p, _p] + [_c_int] * 6 + [_p, _p, _c_f64,
                                                                                 _c_f64, _c_int, _p,
                                                                                 _c_size, _p, _p]),
    "cvqpg_blasm_reconstruct": (_c_int, [_p, _p] + [_c_int] * 4 + [_p, _p, _c_f64, _c_f64, _c_int,
                                                                   _p, _c_size, _p, _p, _p]),
    "cvqpg_blasm_reconstruct_backward": (_c_int, [_p] * 4 + [_c_int] * 4 + [_p, _p, _c_f64, _c_f64,
                                                                            _c_int, _p, _c_size,
                                                                            _p, _p, _p]),
    "cvqpg_blasm_reconstruct_aperture": (_c_int, [_p, _p, _p] + [_c_int] * 4 + [_p, _p, _c_f64,
                                                                             _c_int, _p, _c_size,
                                                                             _p, _p, _p]),
    "cvqpg_blasm_reconstruct_aperture_backward": (_c_int, [_p] * 5 + [_c_int] * 4 +
                                                  [_p, _p, _c_f64, _c_int, _p, _c_size, _p, _p,
                                                   _p]),
    "cvqpg_loss_workspace_size": (_c_size, [_c_int] * 4),
    "cvqpg_loss_forward_backward": (_c_int, [_p, _p, _p, _p, _p, ctypes.c_uint64, _p, _c_int,
                                             _c_int, _c_int, _c_int, _p, _c_size, _p, _p, _p]),
    "cvqpg_loss_noise": (_c_int, [ctypes.c_uint64, _p] + [_c_int] * 4 + [_p, _p]),
    "cvqpg_loss_noalloc_max_bytes": (_c_size, []),
    "cvqpg_loss_forward_backward_noalloc": (_c_int, [_p, _p, _p, _p, _p, _p, ctypes.c_uint64, _p,
                                                     _c_int, _c_int, _c_int, _c_int, _p, _p, _p]),
    "cvqpg_loss_target_stats_size": (_c_size, [_c_int] * 4),
    "cvqpg_loss_target_stats": (_c_int, [_p] + [_c_int] * 4 + [_p, _c_size, _p]),
    "cvqpg_loss_forward_backward_stats": (_c_int, [_p, _p, _p, _p, _p, _p, ctypes.c_uint64, _p,
                                                   _c_int, _c_int, _c_int, _c_int, _p, _c_size,
                                                   _p, _p, _p]),
    "cvqpg_plane_l2_workspace_size": (_c_size, [_c_i64]),
    "cvqpg_plane_l2": (_c_int, [_p, _p, _p, _p, _c_f32, _c_f32, _c_f32, _c_f32, _c_i64, _p, _c_size,
                                _p, _p, _p]),
    "cvqpg_ssim_term": (_c_int, [_p, _p, _c_int, _c_int, _c_int, _p, _c_size, _p, _p, _p]),
    "cvqpg_prologue_forward": (_c_int, [_p, _c_int, _c_int, _c_int, _c_int, _p, _c_f32, _c_f32,
                                        _c_int, _c_f32, _p, _c_size, _c_i64, _p, _p, _p]),
    "cvqpg_prologue_forward_live": (_c_int, [_p, _c_int, _c_int, _c_int, _c_int, _p, _c_f32, _c_f32,
                                        _c_int, _c_f32, _p, _c_size, _c_i64, _p, _p, _p]),
    "cvqpg_bin_scan": (_c_int, [_c_int, _p, _c_size, _c_i64, _p, _p]),
    "cvqpg_raster_backward_raw": (_c_int, [_p, _c_int, _c_int, _c_int, _c_int, _c_f32, _c_f32,
                                           _c_int, _c_f32, _p, _c_int, _p, _c_size, _c_i64, _p,
                                           _p, _p]),
    "cvqpg_adan_step": (_c_int, [_p, _p, _p, _p, _p, _c_int, _p, _c_f32, _c_f32, _c_int, _c_f32,
                                 _c_f32, _c_f32, _c_f32, _c_f32, _c_int, _c_int, _p, _p, _p]),
    "cvqpg_raster_overflow_flags": (_c_int, [_p, _c_size, _c_int, _c_i64,
                                             ctypes.POINTER(ctypes.c_void_p)]),
    "cvqpg_raster_clear_overflow": (_c_int, [_p, _c_size, _c_int, _c_i64, _p]),
    "cvqpg_soft_window_pad": (_c_int, [_p] + [_c_int] * 5 + [_p, _p]),
    "cvqpg_dpac_encode": (_c_int, [_p] + [_c_int] * 3 + [_p, _p]),
    "cvqpg_metrics_workspace_size": (_c_size, [_c_int] * 4),
    "cvqpg_plane_metrics": (_c_int, [_p, _c_int, _c_int, _p] + [_c_int] * 4 + [_p, _c_size, _p,
                                                                             _p, _p]),
    "cvqpg_nn_conv": (_c_int, [_p] + [_c_int] * 4 + [_p, _p] + [_c_int] * 3 + [_p] * 5),
    "cvqpg_nn_wgrad_workspace_size": (_c_size, [_c_int] * 6),
    "cvqpg_nn_conv_wgrad": (_c_int, [_p] + [_c_int] * 4 + [_p, _c_int, _c_int, _p, _p, _p,
                                                          _c_size, _p]),
    "cvqpg_nn_flip_filter": (_c_int, [_p, _c_int, _c_int, _c_int, _p, _p]),
    "cvqpg_nn_relu_backward": (_c_int, [_p, _p, _c_i64, _p, _p]),
    "cvqpg_nn_pool2": (_c_int, [_p] + [_c_int] * 4 + [_p, _p]),
    "cvqpg_nn_pool2_backward": (_c_int, [_p] + [_c_int] * 4 + [_p, _c_int, _p]),
    "cvqpg_nn_upsample2": (_c_int, [_p] + [_c_int] * 6 + [_p, _p]),
    "cvqpg_nn_upsample2_backward": (_c_int, [_p] + [_c_int] * 6 + [_p, _c_int, _p]),
    "cvqpg_nn_tiles_from_image": (_c_int, [_p] + [_c_int] * 4 + [_p, _p]),
    "cvqpg_nn_wmse_workspace_size": (_c_size, [_c_i64]),
    "cvqpg_nn_wmse": (_c_int, [_p, _p, _p, _c_i64, _c_f32, _p, _p, _c_size, _p, _p]),
    "cvqpg_nn_adam": (_c_int, [_p, _p, _p, _p, _p, _c_i64, _c_f32, _c_f32, _c_f32, _c_f32,
                               _c_int, _p]),
    "cvqpg_nn_to_bf16": (_c_int, [_p, _c_i64, _p, _p]),
}

_LIB = None


class CvqpgError(RuntimeError):
    pass


def lib():
    """Load (once) and return the ctypes handle; raises if the library is absent."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise CvqpgError(f"{LIB_PATH} is not built; run build.py (hipcc, gfx950)")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


class RasterScene(ctypes.Structure):
    """cvqpg_raster_scene of include/cvqpg.h."""
    _fields_ = [("n", _c_int), ("width", _c_int), ("height", _c_int), ("channels", _c_int),
                ("ws", _p), ("ws_bytes", _c_size), ("capacity", _c_i64),
                ("point_list", _p), ("ranges", _p), ("field", _p), ("grad_field", _p),
                ("k", _p)]


def loaded_path():
    return LIB_PATH if _LIB is not None else None


def check(rc):
    if rc != 0:
        msg = lib().cvqpg_last_error().decode(errors="replace")
        raise CvqpgError(msg or f"libcvqpg error {rc}")


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def doubles(values):
    """Host array of doubles for the C ABI's small parameter lists."""
    vals = [float(v) for v in values]
    return (ctypes.c_double * len(vals))(*vals)


def ptr_array(tensors):
    """Host array of device pointers (e.g. the seven raw parameter tensors)."""
    return (ctypes.c_void_p * len(tensors))(*[t.data_ptr() for t in tensors])


def int64s(values):
    vals = [int(v) for v in values]
    return (ctypes.c_int64 * len(vals))(*vals)


def floats(values):
    vals = [float(v) for v in values]
    return (ctypes.c_float * len(vals))(*vals)


def stream_handle(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
