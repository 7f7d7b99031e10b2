// One plane's three-term masked L2 and its gradient for gfx950: the call
// multiplane_loss_odak.__call__(image, target, plane_id, inject_noise, noise_ratio)
// (utils/propagator.py:417-449) as a stand-alone entry point, for callers that use the
// reference's loss object rather than the fused multi-plane kernel of loss.hip.
//
// With e1 = I - T, T~ = T + noise * noise_ratio * (max T - min T) (noise NULL: T~ = T),
// e2 = I - T~ and n elements:
//   L     = w0/n sum e1^2 + w1/n sum M^2 e2^2 + w2/n sum T~^2 e2^2
//           (MSE(I, T), MSE(I M, T~ M), MSE(I T~, T~ T~))
//   dL/dI = 2/n (w0 e1 + (w1 M^2 + w2 T~^2) e2)
// Three launches: the target's range (only with noise), the element pass with per-block
// partial sums, and one workgroup that adds the partials in a fixed order, so the result
// is deterministic and nothing is read back to the host.
#include "common.h"
#include "cvqpg.h"

namespace cvqpg {

constexpr int PL_NT = 256;
constexpr int PL_MAXB = 1024;

__device__ __forceinline__ float block_sum(float v, float* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int off = PL_NT / 2; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  const float r = red[0];
  __syncthreads();
  return r;
}

// per-block (min, max) of the target.  grid nb
__global__ __launch_bounds__(PL_NT) void k_target_range(const float* __restrict__ target,
                                                        int64_t n, float2* __restrict__ part) {
  __shared__ float lo[PL_NT], hi[PL_NT];
  const int t = threadIdx.x;
  float a = INFINITY, b = -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * PL_NT + t; i < n; i += (int64_t)gridDim.x * PL_NT) {
    const float v = target[i];
    a = fminf(a, v);
    b = fmaxf(b, v);
  }
  lo[t] = a;
  hi[t] = b;
  __syncthreads();
  for (int off = PL_NT / 2; off > 0; off >>= 1) {
    if (t < off) {
      lo[t] = fminf(lo[t], lo[t + off]);
      hi[t] = fmaxf(hi[t], hi[t + off]);
    }
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = make_float2(lo[0], hi[0]);
}

// element pass.  grid nb
__global__ __launch_bounds__(PL_NT) void k_plane_l2(const float* __restrict__ image,
                                                    const float* __restrict__ target,
                                                    const float* __restrict__ mask,
                                                    const float* __restrict__ noise,
                                                    const float2* __restrict__ range_part,
                                                    float noise_ratio, float w0, float w1,
                                                    float w2, int64_t n,
                                                    float* __restrict__ grad,
                                                    float* __restrict__ part) {
  __shared__ float red[PL_NT];
  __shared__ float s_scale;
  const int t = threadIdx.x;
  if (noise) {  // every workgroup reduces the (at most PL_MAXB) range partials itself
    float a = INFINITY, b = -INFINITY;
    for (int k = t; k < (int)gridDim.x; k += PL_NT) {
      const float2 r = range_part[k];
      a = fminf(a, r.x);
      b = fmaxf(b, r.y);
    }
    red[t] = a;
    __syncthreads();
    for (int off = PL_NT / 2; off > 0; off >>= 1) {
      if (t < off) red[t] = fminf(red[t], red[t + off]);
      __syncthreads();
    }
    const float mn = red[0];
    __syncthreads();
    red[t] = b;
    __syncthreads();
    for (int off = PL_NT / 2; off > 0; off >>= 1) {
      if (t < off) red[t] = fmaxf(red[t], red[t + off]);
      __syncthreads();
    }
    if (t == 0) s_scale = noise_ratio * (red[0] - mn);
    __syncthreads();
  }
  const float scale = noise ? s_scale : 0.f;
  const float g2 = 2.0f / (float)n;
  float acc = 0.f;
  for (int64_t i = (int64_t)blockIdx.x * PL_NT + t; i < n; i += (int64_t)gridDim.x * PL_NT) {
    const float I = image[i], T = target[i], M = mask[i];
    const float Tn = noise ? fmaf(noise[i], scale, T) : T;
    const float e1 = I - T, e2 = I - Tn;
    const float wm = w1 * M * M + w2 * Tn * Tn;
    acc += w0 * e1 * e1 + wm * e2 * e2;
    if (grad) grad[i] = g2 * (w0 * e1 + wm * e2);
  }
  const float s = block_sum(acc, red);
  if (t == 0) part[blockIdx.x] = s;
}

// loss[0] = (1/n) sum of the partials, in a fixed order.  one workgroup
__global__ __launch_bounds__(PL_NT) void k_plane_l2_total(const float* __restrict__ part, int nb,
                                                          int64_t n, float* __restrict__ loss) {
  __shared__ double red[PL_NT];
  const int t = threadIdx.x;
  double a = 0.0;
  for (int k = t; k < nb; k += PL_NT) a += (double)part[k];
  red[t] = a;
  __syncthreads();
  for (int off = PL_NT / 2; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  if (t == 0) loss[0] = (float)(red[0] / (double)n);
}

static int plane_blocks(int64_t n) {
  const int64_t b = (n + PL_NT * 4 - 1) / (PL_NT * 4);
  return (int)(b < 1 ? 1 : (b > PL_MAXB ? PL_MAXB : b));
}

struct PlaneWS {
  float2* range;
  float* part;
  size_t bytes;
};

static PlaneWS carve_plane(void* base) {
  Carver cv(base);
  PlaneWS w;
  w.range = cv.take<float2>(PL_MAXB);
  w.part = cv.take<float>(PL_MAXB);
  w.bytes = cv.bytes();
  return w;
}

}  // namespace cvqpg

using namespace cvqpg;

extern "C" size_t cvqpg_plane_l2_workspace_size(int64_t n) {
  (void)n;
  return carve_plane(nullptr).bytes;
}

extern "C" int cvqpg_plane_l2(const float* image, const float* target, const float* mask,
                              const float* noise, float noise_ratio, float w0, float w1,
                              float w2, int64_t n, void* ws, size_t ws_bytes, float* grad_image,
                              float* loss, void* stream) {
  CVQPG_CHECK_ARG(n > 0, "empty input (n = %lld)", (long long)n);
  CVQPG_CHECK_ARG(image && target && mask && loss && ws, "NULL argument pointer");
  const PlaneWS need = carve_plane(nullptr);
  CVQPG_CHECK_ARG(ws_bytes >= need.bytes, "workspace too small: %zu < %zu", ws_bytes, need.bytes);
  const PlaneWS w = carve_plane(ws);
  hipStream_t s = (hipStream_t)stream;
  const int nb = plane_blocks(n);
  if (noise)
    hipLaunchKernelGGL(k_target_range, dim3(nb), dim3(PL_NT), 0, s, target, n, w.range);
  hipLaunchKernelGGL(k_plane_l2, dim3(nb), dim3(PL_NT), 0, s, image, target, mask, noise,
                     (const float2*)w.range, noise_ratio, w0, w1, w2, n, grad_image, w.part);
  hipLaunchKernelGGL(k_plane_l2_total, dim3(1), dim3(PL_NT), 0, s, (const float*)w.part, nb, n,
                     loss);
  CVQPG_LAUNCH_CHECK();
  return 0;
}
