/*
 * libcvqpg -- C ABI of the MI355X (gfx950) CVQPG hologram hot path.
 *
 * Every entry point:
 *   - takes plain device pointers and sizes (no framework types), and a HIP stream
 *     handle passed as `void *` (a hipStream_t; NULL is the null stream);
 *   - is stream-ordered: it enqueues kernels on `stream` and never synchronises
 *     the host, never allocates device memory (scratch lives in a caller-owned
 *     workspace sized by the matching *_workspace_size function), so it can be
 *     captured into a HIP graph;
 *   - returns 0 on success, 1 on an invalid argument, 2 on a HIP runtime error;
 *     cvqpg_last_error() returns the message of the last failure on the calling
 *     thread.
 *
 * Layouts: complex fields are complex64 interleaved (re, im) with shape [C, H, W];
 * per-Gaussian arrays are row-major [N, k] float32; ids and ranges are int32.
 */
#ifndef CVQPG_H
#define CVQPG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed call on this thread ("" if none). */
const char *cvqpg_last_error(void);

/* ABI version of this header (bumped on any signature change). */
int cvqpg_abi_version(void);

/* ------------------------------------------------------------------------ */
/* Tile binning and rasterisation.                                           */
/* Replaces the native API of cuda/qpf_rasterizer.h:30-57 (qpf_render_forward */
/* and qpf_render_backward, implemented in cuda/raster.cu:30-143) that        */
/* cuda/rasterizer.py:43-55 binds through pybind11 (cuda/binding.cpp:4-6).   */
/* ------------------------------------------------------------------------ */

/* Bytes of workspace for n Gaussians, `channels` (1 or 3), a width x height
 * image and room for `capacity` (tile, Gaussian) keys.  The same workspace must
 * be passed to every raster call of one forward/backward pair: it carries the
 * binning (offsets, rectangles, keys) from the forward to the backward. */
size_t cvqpg_raster_workspace_size(int n, int channels, int width, int height, int64_t capacity);

/* Per-Gaussian 5-sigma tile count and its exclusive prefix sum; writes the total
 * key count P to *num_pairs (device int32).
 * Replaces preprocessKernel + cub::DeviceScan::InclusiveSum
 * (cuda/forward.cu:15-26, cuda/raster.cu:62-69).  Unlike the reference it does
 * not read P back to the host (cuda/raster.cu:71); the caller decides. */
int cvqpg_bin_count(const float *means, const int32_t *radii, int n, int width, int height,
                    void *ws, size_t ws_bytes, int64_t capacity, int32_t *num_pairs,
                    void *stream);

/* Emits the (tile, Gaussian) keys, sorts them by tile with a stable LDS radix
 * sort (ids ascend within a tile, so the result equals the reference's 64-bit
 * key sort bit for bit) and writes point_list [max(P,1)] (sorted Gaussian ids;
 * [0] when P == 0) and ranges [T, 2] ([start, end), (0, 0) for untouched tiles).
 * If P > capacity nothing is sorted, every range is (0, 0) and the caller must
 * retry with a larger workspace.
 * Replaces duplicateWithKeysKernel, cub::DeviceRadixSort::SortPairs and
 * identifyTileRangesKernel (cuda/forward.cu:28-60, cuda/raster.cu:77-95). */
int cvqpg_bin_sort(const float *means, const int32_t *radii, int n, int width, int height,
                   void *ws, size_t ws_bytes, int64_t capacity, int32_t *point_list,
                   int32_t *ranges, void *stream);

/* Forward render: field[c, y, x] = sum over the tile's Gaussians g of
 * colours[g,c] * opacities[g] * G * exp(i (phase[g,c] + k[c] * base)), written
 * as complex64 for every pixel (untouched tiles get 0).
 * Replaces renderTileKernel<C> (cuda/forward.cu:62-162) and the output
 * allocation/interleave of cuda/raster.cu:54-56 and cuda/rasterizer.py:48. */
int cvqpg_render_forward(const float *means, const float *inv_cov, const float *colours,
                         const float *phase, const float *opacities, const float *curvature,
                         const float *area, const float *k, int n, int channels, float dxy2,
                         int width, int height, void *ws, size_t ws_bytes, int64_t capacity,
                         const int32_t *point_list, const int32_t *ranges, float *field,
                         void *stream);

/* Adjoint of cvqpg_render_forward for grad_field = dL/dRe + i dL/dIm (complex64
 * [C,H,W]); writes d_means [N,2], d_inv_cov [N,3], d_colours [N,C],
 * d_phase [N,C], d_opacities [N], d_curvature [N], d_area [N] (overwritten, not
 * accumulated).  Per-tile recomputation, no stored per-pixel state; partial sums
 * are reduced per Gaussian without atomics, so the result is deterministic.
 * Replaces qpf_render_backward + renderBackwardKernel<C>
 * (cuda/raster.cu:107-143, cuda/backward.cu:22-182). */
int cvqpg_render_backward(const float *grad_field, const float *means, const float *inv_cov,
                          const float *colours, const float *phase, const float *opacities,
                          const float *curvature, const float *area, const float *k, int n,
                          int channels, float dxy2, int width, int height, void *ws,
                          size_t ws_bytes, int64_t capacity, const int32_t *point_list,
                          const int32_t *ranges, float *d_means, float *d_inv_cov,
                          float *d_colours, float *d_phase, float *d_opacities,
                          float *d_curvature, float *d_area, void *stream);

/* cvqpg_bin_count + cvqpg_bin_sort + cvqpg_render_forward in one stream-ordered
 * call with no host read-back (graph-capturable); *num_pairs receives P.
 * Replaces qpf_render_forward (cuda/raster.cu:30-105). */
int cvqpg_qpf_render_forward(const float *means, const float *inv_cov, const float *colours,
                             const float *phase, const float *opacities, const float *curvature,
                             const float *area, const int32_t *radii, const float *k, int n,
                             int channels, float dxy2, int width, int height, void *ws,
                             size_t ws_bytes, int64_t capacity, float *field,
                             int32_t *point_list, int32_t *ranges, int32_t *num_pairs,
                             void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CVQPG_H */
