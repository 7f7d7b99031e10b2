// Hologram export and evaluation on gfx950: guard-band padding of the rendered field,
// double-phase amplitude coding (DPAC) to 8-bit phase levels, and per-plane PSNR / SSIM
// of reconstructions against the multi-plane targets.
//
// Reference behaviour:
//   soft_window_pad            model_2d_gaussian_rgb.py:186-241
//   dpac + 8-bit quantisation  export_poh.py:110-119, 173-178
//   PSNR / SSIM per plane      train_2d_gaussian_rgb.py:98-112 (skimage
//                              peak_signal_noise_ratio and structural_similarity with
//                              data_range 1, channel_axis 2: 7x7 uniform window, sample
//                              covariance N/(N-1), K = (0.01, 0.03), mean over the
//                              interior 3 pixels away from the border, mean over the
//                              channels)
#include <algorithm>

#include "common.h"
#include "cvqpg.h"
#include "window.h"

namespace cvqpg {

constexpr float TWO_PI_F = 6.283185307179586f;

// padded[c][r][q] = replicate-or-centre(field) * wy(r) * wx(q).  grid-stride over C pH pW
__global__ void k_soft_window_pad(const float2* __restrict__ field, int C, int H, int W, int pH,
                                  int pW, float2* __restrict__ out) {
  const int pad_h = (pH - H) / 2, pad_w = (pW - W) / 2;
  const int64_t n = (int64_t)C * pH * pW;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int q = (int)(i % pW);
    const int r = (int)((i / pW) % pH);
    const int c = (int)(i / ((int64_t)pW * pH));
    const int sr = min(max(r - pad_h, 0), H - 1), sc = min(max(q - pad_w, 0), W - 1);
    const float w = taper(r, pH, pad_h, H) * taper(q, pW, pad_w, W);
    const float2 h = field[((size_t)c * H + sr) * W + sc];
    out[i] = make_float2(h.x * w, h.y * w);
  }
}

// Python's a % b for b > 0 (result in [0, b))
__device__ __forceinline__ float pymod(float a, float b) {
  float r = fmodf(a, b);
  if (r < 0.0f) r += b;
  return r >= b ? r - b : r;
}

// levels = round(dpac(phase, amp) / 2pi * 256) mod 256 with phase = angle(h) mod 2pi,
// amp = clamp(|h|, 0, 1), dpac = (phase -+ acos(amp)) on the checkerboard, mod 2pi.
// round() is torch.round: half to even.
__global__ void k_dpac(const float2* __restrict__ field, int C, int pH, int pW,
                       uint8_t* __restrict__ levels) {
  const int64_t n = (int64_t)C * pH * pW;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % pW);
    const int y = (int)((i / pW) % pH);
    const float2 h = field[i];
    const float ph = pymod(atan2f(h.y, h.x), TWO_PI_F);
    const float amp = fminf(fmaxf(sqrtf(h.x * h.x + h.y * h.y), 0.0f), 1.0f);
    const float off = acosf(amp);
    // (even, even) and (odd, odd): phase - offset; the other two: phase + offset
    const float v = pymod(((x + y) & 1) ? ph + off : ph - off, TWO_PI_F);
    const float lv = rintf(v / TWO_PI_F * 256.0f);
    levels[i] = (uint8_t)(((int)lv) & 255);
  }
}

// ---------------------------------------------------------------------------------
// metrics: per (plane, channel) image, tiles of 32 x 32 interior SSIM positions;
// each workgroup writes its partial sum of S and of the squared error of its pixels.
// ---------------------------------------------------------------------------------
constexpr int MT = 32;           // SSIM positions per tile side
constexpr int MW = 7;            // window
constexpr int MR = MW / 2;       // 3
constexpr int MI = MT + MW - 1;  // 38 staged pixels per side

struct MetricDims {
  int D, C, H, W, rH, rW, r0, c0;  // recon [D][C][rH][rW], crop origin (r0, c0)
  int tx, ty;                      // tiles per image
};

__device__ __forceinline__ float clamp01f(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

// grid (tx * ty, D * C), 256 threads
__global__ __launch_bounds__(256) void k_metric_partials(const float* __restrict__ recon,
                                                         const float* __restrict__ target,
                                                         MetricDims m,
                                                         float2* __restrict__ part) {
  __shared__ float sx[MI][MI + 1];
  __shared__ float sy[MI][MI + 1];
  __shared__ float hs[5][MI][MT + 1];
  __shared__ float red[2][256];
  const int dc = blockIdx.y;
  const int tile = blockIdx.x;
  const int ox = (tile % m.tx) * MT, oy = (tile / m.tx) * MT;   // first SSIM position - 3
  const int t = threadIdx.x;
  const float* rp = recon + (size_t)dc * m.rH * m.rW;
  const float* tp = target + (size_t)dc * m.H * m.W;
  // stage pixels [oy, oy + 38) x [ox, ox + 38) of the cropped, clamped reconstruction
  float sse = 0.f;
  for (int e = t; e < MI * MI; e += 256) {
    const int yy = e / MI, xx = e % MI;
    const int y = oy + yy, x = ox + xx;
    float a = 0.f, b = 0.f;
    if (y < m.H && x < m.W) {
      a = clamp01f(rp[(size_t)(m.r0 + y) * m.rW + m.c0 + x]);
      b = tp[(size_t)y * m.W + x];
      // each pixel's squared error is counted by the tile that owns it
      if (yy < MT && xx < MT) sse += (a - b) * (a - b);
    }
    sx[yy][xx] = a;
    sy[yy][xx] = b;
  }
  __syncthreads();
  // horizontal 7-sums of x, y, xx, yy, xy
  for (int e = t; e < MI * MT; e += 256) {
    const int yy = e / MT, xx = e % MT;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f;
#pragma unroll
    for (int k = 0; k < MW; ++k) {
      const float a = sx[yy][xx + k], b = sy[yy][xx + k];
      s0 += a;
      s1 += b;
      s2 = fmaf(a, a, s2);
      s3 = fmaf(b, b, s3);
      s4 = fmaf(a, b, s4);
    }
    hs[0][yy][xx] = s0;
    hs[1][yy][xx] = s1;
    hs[2][yy][xx] = s2;
    hs[3][yy][xx] = s3;
    hs[4][yy][xx] = s4;
  }
  __syncthreads();
  const float C1 = 0.0001f, C2 = 0.0009f;           // (0.01 * 1)^2, (0.03 * 1)^2
  const float inv_np = 1.0f / (MW * MW), cov_norm = (float)(MW * MW) / (MW * MW - 1);
  float ssum = 0.f;
  for (int e = t; e < MT * MT; e += 256) {
    const int yy = e / MT, xx = e % MT;
    // SSIM position (oy + yy + 3, ox + xx + 3) must lie in [3, H - 3) x [3, W - 3)
    if (oy + yy + MR >= m.H - MR || ox + xx + MR >= m.W - MR) continue;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f;
#pragma unroll
    for (int k = 0; k < MW; ++k) {
      s0 += hs[0][yy + k][xx];
      s1 += hs[1][yy + k][xx];
      s2 += hs[2][yy + k][xx];
      s3 += hs[3][yy + k][xx];
      s4 += hs[4][yy + k][xx];
    }
    const float ux = s0 * inv_np, uy = s1 * inv_np;
    const float vx = cov_norm * (s2 * inv_np - ux * ux);
    const float vy = cov_norm * (s3 * inv_np - uy * uy);
    const float vxy = cov_norm * (s4 * inv_np - ux * uy);
    const float A1 = 2.f * ux * uy + C1, A2 = 2.f * vxy + C2;
    const float B1 = ux * ux + uy * uy + C1, B2 = vx + vy + C2;
    ssum += (A1 * A2) / (B1 * B2);
  }
  red[0][t] = ssum;
  red[1][t] = sse;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off) {
      red[0][t] += red[0][t + off];
      red[1][t] += red[1][t + off];
    }
    __syncthreads();
  }
  if (t == 0) part[(size_t)dc * gridDim.x + tile] = make_float2(red[0][0], red[1][0]);
}

// psnr[d] = 10 log10(1 / mse_d), ssim[d] = mean over channels of the mean interior S.
// One workgroup per plane; fixed summation order.
__global__ __launch_bounds__(256) void k_metric_final(const float2* __restrict__ part,
                                                      MetricDims m, float* __restrict__ psnr,
                                                      float* __restrict__ ssim) {
  __shared__ double red[2][256];
  const int d = blockIdx.x, t = threadIdx.x;
  const int nt = m.tx * m.ty;
  double s_ssim = 0.0, s_sse = 0.0;
  for (int c = 0; c < m.C; ++c) {
    double sc = 0.0;
    for (int i = t; i < nt; i += 256) {
      const float2 p = part[(size_t)(d * m.C + c) * nt + i];
      sc += p.x;
      s_sse += p.y;
    }
    s_ssim += sc / ((double)(m.H - 2 * MR) * (double)(m.W - 2 * MR));
  }
  red[0][t] = s_ssim;
  red[1][t] = s_sse;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off) {
      red[0][t] += red[0][t + off];
      red[1][t] += red[1][t + off];
    }
    __syncthreads();
  }
  if (t == 0) {
    const double mse = red[1][0] / ((double)m.C * m.H * m.W);
    psnr[d] = (float)(10.0 * log10(1.0 / mse));
    ssim[d] = (float)(red[0][0] / m.C);
  }
}

static MetricDims metric_dims(int D, int C, int H, int W, int rH, int rW) {
  MetricDims m;
  m.D = D; m.C = C; m.H = H; m.W = W; m.rH = rH; m.rW = rW;
  m.r0 = rH / 2 - H / 2;   // odak crop_center
  m.c0 = rW / 2 - W / 2;
  m.tx = ceil_div(W, MT);
  m.ty = ceil_div(H, MT);
  return m;
}

}  // namespace cvqpg

using namespace cvqpg;

extern "C" int cvqpg_soft_window_pad(const float* field, int channels, int height, int width,
                                     int pad_height, int pad_width, float* out, void* stream) {
  CVQPG_CHECK_ARG(field && out, "NULL field/out");
  CVQPG_CHECK_ARG(channels >= 1 && height > 0 && width > 0 && pad_height >= height &&
                      pad_width >= width,
                  "invalid sizes");
  const int64_t n = (int64_t)channels * pad_height * pad_width;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_soft_window_pad, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                     reinterpret_cast<const float2*>(field), channels, height, width, pad_height,
                     pad_width, reinterpret_cast<float2*>(out));
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_dpac_encode(const float* field, int channels, int pad_height, int pad_width,
                                 uint8_t* levels, void* stream) {
  CVQPG_CHECK_ARG(field && levels, "NULL field/levels");
  CVQPG_CHECK_ARG(channels >= 1 && pad_height > 0 && pad_width > 0, "invalid sizes");
  const int64_t n = (int64_t)channels * pad_height * pad_width;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_dpac, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                     reinterpret_cast<const float2*>(field), channels, pad_height, pad_width,
                     levels);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" size_t cvqpg_metrics_workspace_size(int planes, int channels, int height, int width) {
  if (planes < 1 || channels < 1 || height < 1 || width < 1) return 0;
  const MetricDims m = metric_dims(planes, channels, height, width, height, width);
  return (size_t)planes * channels * m.tx * m.ty * sizeof(float2);
}

extern "C" int cvqpg_plane_metrics(const float* recon, int recon_height, int recon_width,
                                   const float* target, int planes, int channels, int height,
                                   int width, void* ws, size_t ws_bytes, float* psnr,
                                   float* ssim, void* stream) {
  CVQPG_CHECK_ARG(recon && target && ws && psnr && ssim, "NULL argument pointer");
  CVQPG_CHECK_ARG(planes >= 1 && channels >= 1, "invalid planes/channels");
  CVQPG_CHECK_ARG(height > 2 * MR && width > 2 * MR, "image smaller than the 7x7 SSIM window");
  CVQPG_CHECK_ARG(recon_height >= height && recon_width >= width,
                  "reconstruction smaller than the target");
  const MetricDims m = metric_dims(planes, channels, height, width, recon_height, recon_width);
  const size_t need = cvqpg_metrics_workspace_size(planes, channels, height, width);
  CVQPG_CHECK_ARG(ws_bytes >= need, "workspace too small: %zu < %zu", ws_bytes, need);
  hipStream_t s = (hipStream_t)stream;
  float2* part = reinterpret_cast<float2*>(ws);
  hipLaunchKernelGGL(k_metric_partials, dim3(m.tx * m.ty, planes * channels), dim3(256), 0, s,
                     recon, target, m, part);
  hipLaunchKernelGGL(k_metric_final, dim3(planes), dim3(256), 0, s, part, m, psnr, ssim);
  CVQPG_LAUNCH_CHECK();
  return 0;
}
