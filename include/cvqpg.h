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

/* cvqpg_bin_count for the raster kernels' own tile lists.  No counterpart in the reference,
 * which keys every tile of the 5-sigma rectangle (cuda/forward.cu:15-26); this counts only
 * the tiles on which the Gaussian's exponent reaches the raster kernels' skip threshold
 * (2^-24 of a unit term) less a small margin, read from the packed records that
 * cvqpg_prologue_forward wrote into ws (call it first).  The forward and adjoint tile
 * kernels skip every other key (the adjoint writes it a zero partial record), so their
 * results are bit for bit those of the full lists; the lists are shorter (about 63 % of
 * the keys at 1920x1080, 150k Gaussians).  Rectangles of more than 64 tiles keep every
 * tile.  The following cvqpg_bin_sort then writes tile lists that are subsequences of the
 * reference's; the kept tiles stay in ws (a flag there marks a live count), where
 * cvqpg_render_backward_tiles and the batched adjoint find each key's partial-record slot. */
int cvqpg_bin_count_live(const float *means, const int32_t *radii, int n, int width,
                         int height, void *ws, size_t ws_bytes, int64_t capacity,
                         int32_t *num_pairs, void *stream);

/* Key-overflow flags of a raster workspace, for callers that size the key buffers once and
 * replay a captured step (the reference reads P back on every call and allocates to fit,
 * cuda/raster.cu:71-80, so it has no counterpart).  Host-only: *flags receives the DEVICE
 * address of two uint32 inside ws: flags[0] = 1 if the last cvqpg_bin_count /
 * cvqpg_qpf_render_forward found more keys than `capacity` (that call then renders nothing
 * and its adjoint is zero), flags[1] = the number of such counts since the last
 * cvqpg_raster_clear_overflow (sticky; no step clears it).  Pass flags as the skip_flag of
 * cvqpg_adan_step so an overflowed step leaves parameters, state and iteration unchanged. */
int cvqpg_raster_overflow_flags(void *ws, size_t ws_bytes, int n, int64_t capacity,
                                uint32_t **flags);
/* Zeroes both flags (and the stored key count) with a kernel on `stream`. */
int cvqpg_raster_clear_overflow(void *ws, size_t ws_bytes, int n, int64_t capacity,
                                void *stream);

/* Emits the (tile, Gaussian) keys, sorts them by tile with a stable LDS radix
 * sort (ids ascend within a tile, so the result equals the reference's 64-bit
 * key sort bit for bit) and writes point_list [max(P,1)] (sorted Gaussian ids;
 * [0] when P == 0) and ranges [T, 2] ([start, end), (0, 0) for untouched tiles).
 * If P > capacity nothing is sorted, every range is (0, 0) and the caller must
 * retry with a larger workspace.  After cvqpg_bin_count_live the keys are the kept
 * tiles only, so each tile's list is an ordered subsequence of the reference's.
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

/* Stages of the two calls above, for callers that reuse the packed parameters
 * between forward and adjoint (the fused training step) or time one kernel:
 *   pack            per-Gaussian derived constants into the workspace
 *                   (the per-pair-invariant part of cuda/forward.cu:98-127);
 *   forward_packed  renderTileKernel<C> on packed parameters (cuda/forward.cu:62-162);
 *   backward_tiles  the per-tile adjoint kernel writing one partial record per key
 *                   (cuda/backward.cu:22-148 without its atomics, :150-165); keys whose
 *                   Gaussian stays below G = 2^-24 on the whole tile are not evaluated
 *                   (their records are zero: each term is below 2^-24 of a unit term);
 *   backward_reduce the per-Gaussian sum of the records and the final gradients. */
int cvqpg_raster_pack(const float *means, const float *inv_cov, const float *colours,
                      const float *phase, const float *opacities, const float *curvature,
                      const float *area, const float *k, int n, int channels, float dxy2,
                      int width, int height, void *ws, size_t ws_bytes, int64_t capacity,
                      void *stream);
int cvqpg_render_forward_packed(int n, int channels, int width, int height, void *ws,
                                size_t ws_bytes, int64_t capacity, const int32_t *point_list,
                                const int32_t *ranges, float *field, void *stream);
int cvqpg_render_backward_tiles(const float *grad_field, const float *k, int n, int channels,
                                int width, int height, void *ws, size_t ws_bytes,
                                int64_t capacity, const int32_t *point_list,
                                const int32_t *ranges, void *stream);
int cvqpg_render_backward_reduce(const float *colours, const float *curvature, const float *area,
                                 int n, int channels, float dxy2, int width, int height, void *ws,
                                 size_t ws_bytes, int64_t capacity, float *d_means,
                                 float *d_inv_cov, float *d_colours, float *d_phase,
                                 float *d_opacities, float *d_curvature, float *d_area,
                                 void *stream);

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

/* Many independent holograms in one launch.  The reference renders one scene per
 * qpf_render call (cuda/rasterizer.py:62-71 -> cuda/raster.cu:30-143); these entry
 * points run the same per-tile forward (cuda/forward.cu:62-162) and adjoint
 * (cuda/backward.cu:22-148) for several scenes from one persistent grid that drains a
 * queue of all scenes' tiles.  Each scene is described as it would be for
 * cvqpg_render_forward_packed / cvqpg_render_backward_tiles (its own workspace, key
 * buffers and field); all scenes share the channel count. */
typedef struct cvqpg_raster_scene {
  int n, width, height, channels;
  void *ws;                 /* raster workspace of this scene (after cvqpg_prologue_forward
                               or cvqpg_raster_pack and cvqpg_bin_count / cvqpg_bin_sort) */
  size_t ws_bytes;
  int64_t capacity;
  const int32_t *point_list, *ranges;
  float *field;             /* forward output, complex64 [C][H][W] */
  const float *grad_field;  /* adjoint input, complex64 [C][H][W] (may be NULL for forward only) */
  const float *k;           /* [C] wave numbers (adjoint) */
} cvqpg_raster_scene;

/* Bytes of the device table for n_scenes scenes (at most 64). */
size_t cvqpg_raster_batch_table_size(int n_scenes);
/* Host-only: validates the scenes and writes the table image into table_host (the
 * caller copies it to a device buffer of the same size, once; the queue counters it
 * holds start at zero and are left at zero by every completed launch).  Scenes are
 * queued in the given order (put the heaviest first). */
int cvqpg_raster_batch_table(const cvqpg_raster_scene *scenes, int n_scenes, void *table_host,
                             size_t table_bytes);
/* Forward / adjoint over every scene of a device table; grid = persistent workgroups,
 * chunk = tiles taken per queue fetch.  Launches that share one table must not overlap. */
int cvqpg_render_forward_batch(void *table, int channels, int grid, int chunk, void *stream);
int cvqpg_render_backward_batch(void *table, int channels, int grid, int chunk, void *stream);

/* ------------------------------------------------------------------------ */
/* Band-limited angular spectrum propagation.                                */
/* The reference has no native code here: these replace the PyTorch/odak    */
/* path soft_window_pad (model_2d_gaussian_rgb.py:186-241) -> reconstruct    */
/* glue (train_2d_gaussian_rgb.py:89-95) -> propagator.reconstruct           */
/* (utils/propagator.py:224-290, custom() :41-64) -> centre crop             */
/* (train_2d_gaussian_rgb.py:193).                                           */
/*                                                                           */
/* Common parameters: `channels` C (1..3) and `planes` D (1..16); the image  */
/* is height x width, the padded field pad_height x pad_width (both even),   */
/* the FFT grid (2 pad_height) x (2 pad_width) (prime factors <= 17, 2       */
/* pad_width <= 8704, 2 pad_height <= 6144).  `wavelengths` [C] and          */
/* `distances` [D] are HOST arrays of doubles (metres); `aperture_radius` is */
/* the circular Fourier aperture radius in samples of the 2x grid (<= 0 or   */
/* larger than the grid: open); `grid` 0 = odak's half-sample frequency grid,*/
/* 1 = exact FFT bins.  The transfer-function phase is evaluated in FP64.    */
/* ------------------------------------------------------------------------ */

/* Workspace bytes; full_plane = 1 for cvqpg_blasm_reconstruct, 0 otherwise. */
size_t cvqpg_blasm_workspace_size(int channels, int planes, int height, int width, int pad_height,
                                  int pad_width, int full_plane);

/* field [C,H,W] complex64 (raster output) -> padded, tapered, amplitude-clamped,
 * propagated to every plane and cropped to H x W: out_field [D,C,H,W] complex64
 * (may be NULL if no backward follows) and intensity [D,C,H,W] float32. */
int cvqpg_blasm_forward(const float *field, int channels, int height, int width, int pad_height,
                        int pad_width, int planes, const double *wavelengths,
                        const double *distances, double pixel_pitch, double aperture_radius,
                        int grid, void *ws, size_t ws_bytes, float *out_field, float *intensity,
                        void *stream);

/* Adjoint of cvqpg_blasm_forward: grad_intensity [D,C,H,W] -> grad_field [C,H,W]
 * complex64 (dL/dRe + i dL/dIm of the raster field), through the crops, the
 * conjugate transfer functions, the zero padding and the amplitude/phase glue. */
int cvqpg_blasm_backward(const float *field, const float *grad_intensity, const float *out_field,
                         int channels, int height, int width, int pad_height, int pad_width,
                         int planes, const double *wavelengths, const double *distances,
                         double pixel_pitch, double aperture_radius, int grid, void *ws,
                         size_t ws_bytes, float *grad_field, void *stream);

/* The two calls above one pass at a time, for callers that time or schedule the passes
 * separately: `passes` is a mask of 1 (first row pass: F1 / A3, which also computes the
 * transform tables in ws), 2 (column pass: F2 / A2) and 4 (last row pass: F3 / A1).  Passes
 * run in that order; a call without pass 1 uses the tables a pass-1 call of the same
 * sizes left in ws.  cvqpg_blasm_forward / _backward are passes = 7.
 * CVQPG_BLASM_TABLES in the mask builds every transform table of these sizes in ws (alone:
 * nothing else runs and the data pointers may be NULL); CVQPG_BLASM_TABLES_KEPT says the
 * tables in ws are current (built by an earlier call with the same sizes), so pass 1 does
 * not rebuild them: a training loop builds them once and saves their launches per step. */
#define CVQPG_BLASM_TABLES 8
#define CVQPG_BLASM_TABLES_KEPT 16
int cvqpg_blasm_forward_passes(int passes, const float *field, int channels, int height,
                               int width, int pad_height, int pad_width, int planes,
                               const double *wavelengths, const double *distances,
                               double pixel_pitch, double aperture_radius, int grid, void *ws,
                               size_t ws_bytes, float *out_field, float *intensity, void *stream);
int cvqpg_blasm_backward_passes(int passes, const float *field, const float *grad_intensity,
                                const float *out_field, int channels, int height, int width,
                                int pad_height, int pad_width, int planes,
                                const double *wavelengths, const double *distances,
                                double pixel_pitch, double aperture_radius, int grid, void *ws,
                                size_t ws_bytes, float *grad_field, void *stream);

/* propagator.reconstruct(phase, amplitude) for one frame per channel
 * (utils/propagator.py:276-290): phase, amplitude [C,pH,pW] float32 ->
 * intensity [D,C,pH,pW] float32 over the whole padded plane, and (if out_field is not
 * NULL) the complex plane fields [D,C,pH,pW] complex64 that the backward needs. */
int cvqpg_blasm_reconstruct(const float *phase, const float *amplitude, int channels,
                            int pad_height, int pad_width, int planes, const double *wavelengths,
                            const double *distances, double pixel_pitch, double aperture_radius,
                            int grid, void *ws, size_t ws_bytes, float *out_field,
                            float *intensity, void *stream);

/* Adjoint of cvqpg_blasm_reconstruct (the autograd backward of propagator.reconstruct with
 * no_grad=False, utils/propagator.py:224-290, as train_2d_gaussian_rgb.py:94 calls it):
 * grad_intensity [D,C,pH,pW] and the forward's out_field -> grad_phase and
 * grad_amplitude [C,pH,pW] float32 (overwritten). */
int cvqpg_blasm_reconstruct_backward(const float *phase, const float *amplitude,
                                     const float *grad_intensity, const float *out_field,
                                     int channels, int pad_height, int pad_width, int planes,
                                     const double *wavelengths, const double *distances,
                                     double pixel_pitch, double aperture_radius, int grid,
                                     void *ws, size_t ws_bytes, float *grad_phase,
                                     float *grad_amplitude, void *stream);

/* cvqpg_blasm_reconstruct / _backward with a custom Fourier aperture instead of the circular
 * one (propagator(aperture=...), utils/propagator.py:195-200 and custom() :41-64): `aperture`
 * is a DEVICE array [2 pad_height][2 pad_width] float32 on the centred frequency grid (the
 * hologram-resolution aperture zero-padded to the 2x grid, as the reference's set_aperture
 * builds it); custom() applies it to the spectrum and to the transfer function, i.e. a
 * factor aperture^2 on every plane. */
int cvqpg_blasm_reconstruct_aperture(const float *phase, const float *amplitude,
                                     const float *aperture, int channels, int pad_height,
                                     int pad_width, int planes, const double *wavelengths,
                                     const double *distances, double pixel_pitch, int grid,
                                     void *ws, size_t ws_bytes, float *out_field,
                                     float *intensity, void *stream);
int cvqpg_blasm_reconstruct_aperture_backward(const float *phase, const float *amplitude,
                                              const float *aperture,
                                              const float *grad_intensity,
                                              const float *out_field, int channels,
                                              int pad_height, int pad_width, int planes,
                                              const double *wavelengths,
                                              const double *distances, double pixel_pitch,
                                              int grid, void *ws, size_t ws_bytes,
                                              float *grad_phase, float *grad_amplitude,
                                              void *stream);

/* ------------------------------------------------------------------------ */
/* Multi-plane loss.  Replaces the PyTorch loss assembly of                  */
/* train_2d_gaussian_rgb.py:191-197: multiplane_loss_odak.__call__           */
/* (utils/propagator.py:417-449, weights [1,1,1]) + GaussianLoss             */
/* (utils/data_utils.py:22-41) per plane, and their autograd backward.       */
/* ------------------------------------------------------------------------ */
size_t cvqpg_loss_workspace_size(int channels, int planes, int height, int width);

/* intensity, target, mask [D,C,H,W] float32 (intensity before the clamp to [0,1]).
 * Target noise for plane d is noise[...] * noise_scale[d] when `noise` is given,
 * otherwise a Philox4x32-10 normal draw keyed by (seed, *step_counter, plane, pixel);
 * noise_scale (HOST [D], 1e-3 * (max T_d - min T_d)) NULL disables it.  Writes
 * grad_intensity [D,C,H,W] = dL/d(intensity) and loss[0] (device) = the summed loss. */
int cvqpg_loss_forward_backward(const float *intensity, const float *target, const float *mask,
                                const float *noise, const float *noise_scale, uint64_t seed,
                                const int32_t *step_counter, int channels, int planes,
                                int height, int width, void *ws, size_t ws_bytes,
                                float *grad_intensity, float *loss, void *stream);

/* The same loss with the target's SSIM window statistics precomputed.  The reference
 * recomputes them every call (GaussianLoss builds a new SSIM module per call,
 * utils/data_utils.py:39-40) although the targets of a fit never change
 * (set_targets, utils/propagator.py:361-415).  cvqpg_loss_target_stats writes
 * (mu_T, E[T^2]) of every valid 11x11 window, [D,C,H-10,W-10] float2
 * (cvqpg_loss_target_stats_size bytes), once per fit;
 * cvqpg_loss_forward_backward_stats then filters only the intensity's moments per call
 * and gives the same gradient, bit for bit, as cvqpg_loss_forward_backward, and the same
 * loss up to the order in which its per-workgroup partial sums are added (target_stats
 * NULL: identical to it). */
/* The unit normals the loss draws in-kernel when `noise` is NULL (Philox4x32-10 keyed by
 * (seed, *step_counter, plane, pixel)), written as noise [D,C,H,W] float32.  Passed back
 * as `noise` with the same noise_scale, they give the same loss and gradient bit for bit,
 * so a caller can draw them ahead of the loss (or keep them for a replay).  No counterpart
 * in the reference, whose multiplane_loss_odak draws torch.randn inside the loss
 * (utils/propagator.py:417-449). */
int cvqpg_loss_noise(uint64_t seed, const int32_t *step_counter, int channels, int planes,
                     int height, int width, float *noise, void *stream);

/* No counterpart in the reference: the loss of cvqpg_loss_forward_backward_stats (same
 * kernels, same results bit for bit) for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  It takes no
 * workspace: its scratch (the per-workgroup partial sums) is a static device array of
 * cvqpg_loss_noalloc_max_bytes() bytes (256 KiB) per GPU, held until the process ends;
 * sizes whose scratch exceeds it are rejected (return 1).  It calls no hipMalloc/hipFree.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other threads) race.  Each launch writes every partial it reads, so nothing
 * carries over from one call to the next. */
size_t cvqpg_loss_noalloc_max_bytes(void);
int cvqpg_loss_forward_backward_noalloc(const float *intensity, const float *target,
                                        const float *target_stats, const float *mask,
                                        const float *noise, const float *noise_scale,
                                        uint64_t seed, const int32_t *step_counter,
                                        int channels, int planes, int height, int width,
                                        float *grad_intensity, float *loss, void *stream);

size_t cvqpg_loss_target_stats_size(int channels, int planes, int height, int width);
int cvqpg_loss_target_stats(const float *target, int channels, int planes, int height,
                            int width, float *stats, size_t stats_bytes, void *stream);
int cvqpg_loss_forward_backward_stats(const float *intensity, const float *target,
                                      const float *target_stats, const float *mask,
                                      const float *noise, const float *noise_scale,
                                      uint64_t seed, const int32_t *step_counter, int channels,
                                      int planes, int height, int width, void *ws,
                                      size_t ws_bytes, float *grad_intensity, float *loss,
                                      void *stream);

/* The reference's loss objects as stand-alone calls (the fused step above runs all
 * planes and both terms in one pass):
 * cvqpg_plane_l2 replaces multiplane_loss_odak.__call__(image, target, plane_id,
 * inject_noise, noise_ratio) (utils/propagator.py:417-449): n elements of image, target
 * and mask (mask = masks[plane_id]); noise (NULL: inject_noise False) is the standard
 * normal draw, scaled in-kernel by noise_ratio * (max target - min target); weights
 * (w0, w1, w2) of MSE(I, T), MSE(I M, T~ M), MSE(I T~, T~ T~).  Writes loss[0] (device)
 * and, if grad_image is not NULL, dL/dimage [n].  Workspace: cvqpg_plane_l2_workspace_size
 * (a few KB, independent of n). */
size_t cvqpg_plane_l2_workspace_size(int64_t n);
int cvqpg_plane_l2(const float *image, const float *target, const float *mask,
                   const float *noise, float noise_ratio, float w0, float w1, float w2,
                   int64_t n, void *ws, size_t ws_bytes, float *grad_image, float *loss,
                   void *stream);
/* cvqpg_ssim_term replaces GaussianLoss(pred, target) (utils/data_utils.py:22-41, with its
 * default lambda_ssim 0.025): loss[0] = 0.025 * 0.2 * (1 - SSIM(image, target)) for
 * image, target [C,H,W] (image clamped to [0, 1], its gradient zero outside), and
 * grad_image [C,H,W] = dloss/dimage.  Workspace: cvqpg_loss_workspace_size(C, 1, H, W). */
int cvqpg_ssim_term(const float *image, const float *target, int channels, int height,
                    int width, void *ws, size_t ws_bytes, float *grad_image, float *loss,
                    void *stream);

/* ------------------------------------------------------------------------ */
/* Fused per-Gaussian prologue and its adjoint (the reference runs these in  */
/* PyTorch: Gaussians2D.apply_activations / compute_2d_covariance_elements / */
/* invert_cov_2D, model_2d_gaussian_rgb.py:83-153, and Scene2D._render_cuda  */
/* :428-436).  raw_params: 7 device pointers in the order means_2d [N,2],    */
/* pre_act_scales [N,2], pre_act_rotation [N], colours [N,C], pre_act_phase  */
/* [N,C], pre_act_opacities [N], curvature [N].  curv_mode 0 off, 1 constant */
/* (200 tanh), 2 scale-aware (curv_nyquist lambda / (6 sigma dxy2)).          */
/* ------------------------------------------------------------------------ */

/* Activations, covariance, inverse, area and 5-sigma radius; writes means_out [N,2]
 * and radii_out [N] (for cvqpg_bin_count) and the packed raster record into the raster
 * workspace, so cvqpg_render_forward_packed / _backward_tiles can follow directly. */
int cvqpg_prologue_forward(const float *const *raw_params, int n, int channels, int width,
                           int height, const float *k, float dxy2, float curv_lambda,
                           int curv_mode, float curv_nyquist, void *ws, size_t ws_bytes,
                           int64_t capacity, float *means_out, int32_t *radii_out,
                           void *stream);

/* cvqpg_prologue_forward that also does the live count of cvqpg_bin_count_live (from the
 * values the prologue holds in registers; the same tiles, bit for bit), the prologue fused
 * into the bin-count kernel.  No counterpart in the reference (its activations, radii and
 * tile counts are separate torch launches and preprocessKernel, model_2d_gaussian_rgb.py:
 * 83-153,424-439, cuda/forward.cu:15-26).  Follow it with cvqpg_bin_scan, then
 * cvqpg_bin_sort. */
int cvqpg_prologue_forward_live(const float *const *raw_params, int n, int channels, int width,
                                int height, const float *k, float dxy2, float curv_lambda,
                                int curv_mode, float curv_nyquist, void *ws, size_t ws_bytes,
                                int64_t capacity, float *means_out, int32_t *radii_out,
                                void *stream);

/* The exclusive scan of the per-Gaussian tile counts already in ws (written by
 * cvqpg_prologue_forward_live) and the key total: the second half of cvqpg_bin_count(_live),
 * the reference's cub::DeviceScan::InclusiveSum (cuda/raster.cu:63-69); *num_pairs and the
 * overflow flags as there. */
int cvqpg_bin_scan(int n, void *ws, size_t ws_bytes, int64_t capacity, int32_t *num_pairs,
                   void *stream);

/* After cvqpg_render_backward_tiles: per-Gaussian sum of the partial records, adjoint of
 * the prologue, nan_to_num (train_2d_gaussian_rgb.py:206-208); writes the raw-parameter
 * gradients flat [(7 + 2C) N] in the order above and one sum of squares per 256
 * Gaussians (sumsq_partials [ceil(N/256)]).  The curvature gradient is 0 while
 * *iteration < curv_warmup (frozen parameter). */
int cvqpg_raster_backward_raw(const float *const *raw_params, int n, int channels, int width,
                              int height, float dxy2, float curv_lambda, int curv_mode,
                              float curv_nyquist, const int32_t *iteration, int curv_warmup,
                              void *ws, size_t ws_bytes, int64_t capacity, float *grads,
                              float *sumsq_partials, void *stream);

/* ------------------------------------------------------------------------ */
/* Optimiser tail: clip_grad_norm_(max_norm) + Adan.step + scheduler.step   */
/* (train_2d_gaussian_rgb.py:205-211, utils/optimizer.py:105-262,            */
/* train_2d_gaussian_rgb.py:58-86).  params: the 7 raw tensors (updated in   */
/* place), sizes[7] their element counts, grads flat as produced above,      */
/* state flat [4][total] (exp_avg, exp_avg_sq, exp_avg_diff, neg_pre_grad,   */
/* zero-initialised), group_lr HOST [7] per tensor (entry 0 is replaced by   */
/* the cosine schedule lr0 -> eta_min over num_itrs).  *iteration (device)   */
/* is the 0-based iteration and is incremented at the end.  skip_flag        */
/* (device, may be NULL): when *skip_flag != 0 nothing is updated and the    */
/* iteration is not advanced (pass cvqpg_raster_overflow_flags' pointer).    */
/* ------------------------------------------------------------------------ */
int cvqpg_adan_step(float *const *params, const int64_t *sizes, const float *grads, float *state,
                    const float *sumsq_partials, int n_partials, const float *group_lr,
                    float lr0, float eta_min, int num_itrs, float beta1, float beta2,
                    float beta3, float eps, float max_norm, int curv_warmup, int curv_enabled,
                    int32_t *iteration, const uint32_t *skip_flag, void *stream);

/* ------------------------------------------------------------------------ */
/* Export and evaluation.                                                   */
/* ------------------------------------------------------------------------ */
/* soft_window_pad (model_2d_gaussian_rgb.py:186-241) of a complex64 field [C][H][W] into
 * out [C][pad_height][pad_width]: replicate padding times the Tukey guard-band window. */
int cvqpg_soft_window_pad(const float *field, int channels, int height, int width,
                          int pad_height, int pad_width, float *out, void *stream);

/* DPAC phase-only encoding to 8-bit levels, replacing export_poh.py:110-119,173-178:
 * phase = angle(h) mod 2pi, amplitude = clamp(|h|, 0, 1), phase -+ arccos(amplitude) on
 * the checkerboard, mod 2pi, level = round(. / 2pi * 256) mod 256 (half to even).
 * field complex64 [C][pad_height][pad_width]; levels uint8 of the same shape. */
int cvqpg_dpac_encode(const float *field, int channels, int pad_height, int pad_width,
                      uint8_t *levels, void *stream);

/* Per-plane PSNR and SSIM of reconstructions against the targets, replacing the skimage
 * calls of train_2d_gaussian_rgb.py:98-112 (data_range 1; SSIM: 7x7 uniform window,
 * sample covariance, K = (0.01, 0.03), mean over the interior and over the channels).
 * recon [planes][channels][recon_height][recon_width] (clamped to [0, 1] and centre-
 * cropped to height x width as odak crop_center does), target [planes][channels][height]
 * [width]; psnr and ssim are device arrays [planes].  Workspace from
 * cvqpg_metrics_workspace_size. */
size_t cvqpg_metrics_workspace_size(int planes, int channels, int height, int width);
int cvqpg_plane_metrics(const float *recon, int recon_height, int recon_width,
                        const float *target, int planes, int channels, int height, int width,
                        void *ws, size_t ws_bytes, float *psnr, float *ssim, void *stream);

/* ------------------------------------------------------------------------ */
/* Learned warm start (no counterpart in the reference, whose initialisation */
/* is random, model_2d_gaussian_rgb.py:40-80): BF16 MFMA training and        */
/* inference kernels of a small convolutional encoder-decoder.               */
/* ------------------------------------------------------------------------ */
/* Activations are NHWC bfloat16 ([batch][height][width][channels], raw uint16 bits);
 * filters are [cout][ksize*ksize*cin] bfloat16, biases FP32.  Channel counts are
 * multiples of 16; ksize is 1 or 3 ("same" zero padding, stride 1).
 * cvqpg_nn_conv: y = act(conv(x, w) + bias); flags bit 0 = ReLU, bit 1 = accumulate
 * into y_f32.  Outputs (any of): y_bf16; y2_bf16 = bf16(y_bf16 + addend); y_f32. */
int cvqpg_nn_conv(const uint16_t *x, int batch, int height, int width, int cin,
                  const uint16_t *w, const float *bias, int cout, int ksize, int flags,
                  const uint16_t *addend, uint16_t *y_bf16, uint16_t *y2_bf16, float *y_f32,
                  void *stream);
/* dw[n][tap*cin + c] = sum over pixels of dy[p][n] * x[p + offset(tap)][c] and
 * db[n] = sum over pixels of dy[p][n] (db may be NULL); FP32, deterministic (per-chunk
 * partials in the workspace, added in a fixed order). */
size_t cvqpg_nn_wgrad_workspace_size(int batch, int height, int width, int cin, int cout,
                                     int ksize);
int cvqpg_nn_conv_wgrad(const uint16_t *x, int batch, int height, int width, int cin,
                        const uint16_t *dy, int cout, int ksize, float *dw, float *db, void *ws,
                        size_t ws_bytes, void *stream);
/* Filter of the data gradient: w_flipped[c][tap][n] = w[n][ksize^2 - 1 - tap][c]. */
int cvqpg_nn_flip_filter(const uint16_t *w, int cout, int cin, int ksize, uint16_t *w_flipped,
                         void *stream);
/* dpre = bf16(post > 0 ? dpost : 0) (post NULL: bf16(dpost)). */
int cvqpg_nn_relu_backward(const float *dpost, const uint16_t *post, int64_t n, uint16_t *dpre,
                           void *stream);
/* 2x2 average pooling to ceil(h/2) x ceil(w/2) (mean of the covered pixels) and its
 * adjoint; nearest-neighbour 2x upsampling from ceil(h/2) x ceil(w/2) to h x w and its
 * adjoint.  accumulate = 1 adds into dx. */
int cvqpg_nn_pool2(const uint16_t *x, int batch, int height, int width, int channels,
                   uint16_t *y, void *stream);
int cvqpg_nn_pool2_backward(const float *dy, int batch, int height, int width, int channels,
                            float *dx, int accumulate, void *stream);
int cvqpg_nn_upsample2(const uint16_t *x, int batch, int src_height, int src_width, int channels,
                       int height, int width, uint16_t *y, void *stream);
int cvqpg_nn_upsample2_backward(const float *dy, int batch, int src_height, int src_width,
                                int channels, int height, int width, float *dx, int accumulate,
                                void *stream);
/* img [batch][channels][height][width] FP32 -> one row per 16x16 tile,
 * tiles [batch][ceil(h/16)][ceil(w/16)][channels*256] bf16, element (c*16 + dy)*16 + dx. */
int cvqpg_nn_tiles_from_image(const float *img, int batch, int channels, int height, int width,
                              uint16_t *tiles, void *stream);
/* loss = sum scale * weight * (out - target)^2 (weight NULL = 1), dout = its gradient. */
size_t cvqpg_nn_wmse_workspace_size(int64_t n);
int cvqpg_nn_wmse(const float *out, const float *target, const float *weight, int64_t n,
                  float scale, float *dout, void *ws, size_t ws_bytes, float *loss, void *stream);
/* Adam (bias-corrected, step >= 1) on FP32 master weights, refreshing the bf16 copy. */
int cvqpg_nn_adam(float *w, uint16_t *w_bf16, const float *grad, float *m, float *v, int64_t n,
                  float lr, float beta1, float beta2, float eps, int step, void *stream);
int cvqpg_nn_to_bf16(const float *src, int64_t n, uint16_t *dst, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CVQPG_H */
