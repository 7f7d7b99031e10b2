// Kernels of the learned warm start: a small convolutional encoder-decoder that maps a
// target image and depth map to initial Gaussians per 16x16 tile.  The reference has no
// learned part; its plain initialisation is model_2d_gaussian_rgb.py:40-80 and the fit
// train_2d_gaussian_rgb.py:187-211.  These kernels train and run the network on gfx950.
//
// Layout: activations are NHWC, one row of channels per pixel ("pixels x channels"),
// stored as bfloat16; every convolution is an implicit GEMM
//     Y[p][n] = sum_{tap, c} X[p + offset(tap)][c] * Wt[n][tap * Cin + c]
// on v_mfma_f32_16x16x16_bf16 (bf16 operands, FP32 accumulation), 3x3 "same" (zero
// padding) or 1x1, stride 1.  Down- and up-sampling are separate 2x2 average pooling and
// nearest-neighbour kernels, so the data gradient of every convolution is again a
// stride-1 convolution (with the flipped, transposed filter) and runs on the same kernel.
// The weight gradient is a GEMM over the pixels, one wave per 16x16 block of the filter
// and per 1024-pixel chunk, the chunks added in a fixed order by a second kernel.
//
// MFMA operand layout (16x16x16, 64 lanes; r = lane % 16, g = lane / 16):
//   A (16 x 16, rows x k): lane holds A[r][4g + j], j = 0..3
//   B (16 x 16, k x cols): lane holds B[4g + j][r]
//   D (16 x 16):           lane holds D[4g + j][r]
#include "common.h"
#include "cvqpg.h"

namespace cvqpg {

typedef uint16_t bf16_t;
typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

// round to nearest even (finite inputs)
__device__ __forceinline__ bf16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f4v mfma16(s4v a, s4v b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

enum { CONV_RELU = 1, CONV_ACCUMULATE = 2 };

// One wave: 16 pixels x up to 64 output channels.  Four waves per workgroup along the
// pixels; gridDim.y walks the output channels in blocks of 64.  No LDS, no barriers.
template <int KS>
__global__ __launch_bounds__(256) void k_conv(const bf16_t* __restrict__ X, int B, int H, int W,
                                              int Cin, const bf16_t* __restrict__ Wt,
                                              const float* __restrict__ bias, int Cout, int flags,
                                              const bf16_t* __restrict__ addend,
                                              bf16_t* __restrict__ Yb, bf16_t* __restrict__ Yb2,
                                              float* __restrict__ Yf) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int M = B * H * W;
  const int m0 = (blockIdx.x * 4 + wave) * 16;
  const int n0 = blockIdx.y * 64;
  if (m0 >= M) return;
  const int K = KS * KS * Cin;
  const int r = lane & 15, g = lane >> 4;
  const int nblk = min(4, (Cout - n0) >> 4);
  const int p = m0 + r;
  const bool pv = p < M;
  int b = 0, y = 0, x = 0;
  if (pv) {
    x = p % W;
    const int t = p / W;
    y = t % H;
    b = t / H;
  }
  f4v acc[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) acc[i] = f4v{0.f, 0.f, 0.f, 0.f};
  for (int tap = 0; tap < KS * KS; ++tap) {
    const int ys = y + tap / KS - KS / 2, xs = x + tap % KS - KS / 2;
    const bool v = pv && ys >= 0 && ys < H && xs >= 0 && xs < W;
    const size_t xoff = v ? ((size_t)(b * H + ys) * W + xs) * Cin + 4 * g : 0;
    const bf16_t* wrow = Wt + (size_t)tap * Cin + 4 * g;
    for (int c0 = 0; c0 < Cin; c0 += 16) {
      const s4v a = v ? *reinterpret_cast<const s4v*>(X + xoff + c0) : s4v{0, 0, 0, 0};
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        if (nb < nblk) {
          const s4v w = *reinterpret_cast<const s4v*>(wrow + (size_t)(n0 + nb * 16 + r) * K + c0);
          acc[nb] = mfma16(a, w, acc[nb]);
        }
      }
    }
  }
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    if (nb >= nblk) continue;
    const int n = n0 + nb * 16 + r;
    const float bn = bias ? bias[n] : 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int pix = m0 + 4 * g + j;
      if (pix >= M) continue;
      float v = acc[nb][j] + bn;
      if (flags & CONV_RELU) v = fmaxf(v, 0.f);
      const size_t idx = (size_t)pix * Cout + n;
      if (Yb) {
        const bf16_t h = to_bf16(v);
        Yb[idx] = h;
        if (Yb2) Yb2[idx] = to_bf16(from_bf16(h) + from_bf16(addend[idx]));
      }
      if (Yf) Yf[idx] = (flags & CONV_ACCUMULATE) ? Yf[idx] + v : v;
    }
  }
}

// dW[n][tap * Cin + c] = sum_p dY[p][n] * X[p + offset(tap)][c] and db[n] = sum_p dY[p][n].
// One wave per 16x16 filter block and per chunk of WG_PIX pixels (gridDim.y chunks); each
// writes its partial block to P[chunk][n][k] (and, for the k = 0 blocks, four partial
// bias sums per channel to Q[chunk][g][n]); k_wgrad_reduce adds the chunks in a fixed
// order, so the result is deterministic.
constexpr int WG_PIX = 1024;

template <int KS>
__global__ __launch_bounds__(256) void k_conv_wgrad(const bf16_t* __restrict__ X, int B, int H,
                                                    int W, int Cin, const bf16_t* __restrict__ dY,
                                                    int Cout, float* __restrict__ P,
                                                    float* __restrict__ Q) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int K = KS * KS * Cin;
  const int KB = K >> 4;
  const int blk = blockIdx.x * 4 + wave;
  if (blk >= (Cout >> 4) * KB) return;
  const int n0 = (blk / KB) * 16, k0 = (blk % KB) * 16;
  const int tap = k0 / Cin, c0 = k0 % Cin;
  const int dy = tap / KS - KS / 2, dx = tap % KS - KS / 2;
  const int r = lane & 15, g = lane >> 4;
  const int M = B * H * W;
  const int pbeg = blockIdx.y * WG_PIX, pend = min(M, pbeg + WG_PIX);
  f4v acc = f4v{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  for (int p0 = pbeg; p0 < pend; p0 += 16) {
    // pixel p0 + 4g and its three successors, decoded once and then stepped
    int p = p0 + 4 * g;
    int x = p % W, t = p / W;
    int y = t % H, b = t / H;
    s4v a, w;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      short av = 0, xv = 0;
      if (p < pend) {
        av = (short)dY[(size_t)p * Cout + n0 + r];
        const int ys = y + dy, xs = x + dx;
        if (ys >= 0 && ys < H && xs >= 0 && xs < W)
          xv = (short)X[((size_t)(b * H + ys) * W + xs) * Cin + c0 + r];
      }
      a[j] = av;
      w[j] = xv;
      bsum += from_bf16((bf16_t)av);
      ++p;
      if (++x == W) {
        x = 0;
        if (++y == H) {
          y = 0;
          ++b;
        }
      }
    }
    acc = mfma16(a, w, acc);
  }
  float* out = P + (size_t)blockIdx.y * Cout * K;
#pragma unroll
  for (int j = 0; j < 4; ++j) out[(size_t)(n0 + 4 * g + j) * K + k0 + r] = acc[j];
  if (k0 == 0) Q[((size_t)blockIdx.y * 4 + g) * Cout + n0 + r] = bsum;
}

__global__ void k_wgrad_reduce(const float* __restrict__ P, const float* __restrict__ Q,
                               int chunks, int64_t nw, int Cout, float* __restrict__ dW,
                               float* __restrict__ db) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nw) {
    float s = 0.f;
    for (int c = 0; c < chunks; ++c) s += P[(size_t)c * nw + i];
    dW[i] = s;
  }
  if (db && i < Cout) {
    float s = 0.f;
    for (int c = 0; c < 4 * chunks; ++c) s += Q[(size_t)c * Cout + i];
    db[i] = s;
  }
}

// Wf[c][tap'][n] = Wt[n][KS*KS - 1 - tap'][c]: the filter of the data gradient
__global__ void k_flip(const bf16_t* __restrict__ Wt, int Cout, int Cin, int KK,
                       bf16_t* __restrict__ Wf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)Cout * KK * Cin;
  if (i >= total) return;
  const int n = (int)(i % Cout);
  const int64_t t = i / Cout;
  const int tp = (int)(t % KK);
  const int c = (int)(t / KK);
  Wf[i] = Wt[((size_t)n * KK + (KK - 1 - tp)) * Cin + c];
}

__global__ void k_relu_bwd(const float* __restrict__ dpost, const bf16_t* __restrict__ post,
                           int64_t n, bf16_t* __restrict__ dpre) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v = dpost[i];
  if (post && from_bf16(post[i]) <= 0.f) v = 0.f;
  dpre[i] = to_bf16(v);
}

__global__ void k_pool2(const bf16_t* __restrict__ X, int B, int H, int W, int C,
                        bf16_t* __restrict__ Y) {
  const int Ho = (H + 1) / 2, Wo = (W + 1) / 2;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * Ho * Wo * C) return;
  const int c = (int)(i % C);
  const int64_t q = i / C;
  const int xo = (int)(q % Wo), yo = (int)((q / Wo) % Ho), b = (int)(q / ((int64_t)Wo * Ho));
  float s = 0.f;
  int cnt = 0;
  for (int dy = 0; dy < 2; ++dy)
    for (int dx = 0; dx < 2; ++dx) {
      const int y = 2 * yo + dy, x = 2 * xo + dx;
      if (y < H && x < W) {
        s += from_bf16(X[((size_t)(b * H + y) * W + x) * C + c]);
        ++cnt;
      }
    }
  Y[i] = to_bf16(s / (float)cnt);
}

__global__ void k_pool2_bwd(const float* __restrict__ dY, int B, int H, int W, int C,
                            float* __restrict__ dX, int accumulate) {
  const int Ho = (H + 1) / 2, Wo = (W + 1) / 2;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * H * W * C) return;
  const int c = (int)(i % C);
  const int64_t q = i / C;
  const int x = (int)(q % W), y = (int)((q / W) % H), b = (int)(q / ((int64_t)W * H));
  const int yo = y >> 1, xo = x >> 1;
  const int cnt = (min(2, H - 2 * yo)) * (min(2, W - 2 * xo));
  const float v = dY[((size_t)(b * Ho + yo) * Wo + xo) * C + c] / (float)cnt;
  dX[i] = accumulate ? dX[i] + v : v;
}

__global__ void k_up2(const bf16_t* __restrict__ X, int B, int Hs, int Ws, int C, int Ht, int Wt,
                      bf16_t* __restrict__ Y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * Ht * Wt * C) return;
  const int c = (int)(i % C);
  const int64_t q = i / C;
  const int x = (int)(q % Wt), y = (int)((q / Wt) % Ht), b = (int)(q / ((int64_t)Wt * Ht));
  Y[i] = X[((size_t)(b * Hs + (y >> 1)) * Ws + (x >> 1)) * C + c];
}

__global__ void k_up2_bwd(const float* __restrict__ dY, int B, int Hs, int Ws, int C, int Ht,
                          int Wt, float* __restrict__ dX, int accumulate) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * Hs * Ws * C) return;
  const int c = (int)(i % C);
  const int64_t q = i / C;
  const int xs = (int)(q % Ws), ys = (int)((q / Ws) % Hs), b = (int)(q / ((int64_t)Ws * Hs));
  float s = 0.f;
  for (int dy = 0; dy < 2; ++dy)
    for (int dx = 0; dx < 2; ++dx) {
      const int y = 2 * ys + dy, x = 2 * xs + dx;
      if (y < Ht && x < Wt) s += dY[((size_t)(b * Ht + y) * Wt + x) * C + c];
    }
  dX[i] = accumulate ? dX[i] + s : s;
}

// planes [B][C][H][W] (FP32) -> tiles [B][Ty][Tx][C * 256] (bf16), channel (c*16 + dy)*16 + dx
__global__ void k_tiles_from_image(const float* __restrict__ img, int B, int C, int H, int W,
                                   int Ty, int Tx, bf16_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int CK = C * 256;
  if (i >= (int64_t)B * Ty * Tx * CK) return;
  const int k = (int)(i % CK);
  const int64_t q = i / CK;
  const int tx = (int)(q % Tx), ty = (int)((q / Tx) % Ty), b = (int)(q / ((int64_t)Tx * Ty));
  const int c = k >> 8, dy = (k >> 4) & 15, dx = k & 15;
  const int y = ty * 16 + dy, x = tx * 16 + dx;
  const float v = (y < H && x < W) ? img[(((size_t)b * C + c) * H + y) * W + x] : 0.f;
  out[i] = to_bf16(v);
}

// weighted squared error: dout = 2 * scale * w * (out - target); per-block partial sums
// of scale * w * (out - target)^2, reduced in a fixed order by k_sum_partials
__global__ __launch_bounds__(256) void k_wmse(const float* __restrict__ out,
                                              const float* __restrict__ target,
                                              const float* __restrict__ weight, int64_t n,
                                              float scale, float* __restrict__ dout,
                                              float* __restrict__ partials) {
  __shared__ float red[256];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  float e = 0.f;
  if (i < n) {
    const float d = out[i] - target[i];
    const float w = weight ? weight[i] : 1.f;
    dout[i] = 2.f * scale * w * d;
    e = scale * w * d * d;
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_sum_partials(const float* __restrict__ partials, int n,
                                                      float* __restrict__ total) {
  __shared__ float red[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) s += partials[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = red[0];
}

// Adam on the FP32 master weights; the bf16 copy used by the convolutions is refreshed
// in the same pass.  step is the 1-based update count.
__global__ void k_adam(float* __restrict__ w, bf16_t* __restrict__ wb, const float* __restrict__ grad,
                       float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = grad[i];
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = to_bf16(wi);
}

__global__ void k_to_bf16(const float* __restrict__ src, int64_t n, bf16_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = to_bf16(src[i]);
}

static int blocks_for(int64_t n) { return ceil_div(n > 0 ? n : 1, 256); }

static int check_conv(int batch, int height, int width, int cin, int cout, int ksize) {
  CVQPG_CHECK_ARG(batch > 0 && height > 0 && width > 0, "invalid activation shape %dx%dx%d",
                  batch, height, width);
  CVQPG_CHECK_ARG((int64_t)batch * height * width < ((int64_t)1 << 31), "too many pixels");
  CVQPG_CHECK_ARG(cin > 0 && cin % 16 == 0, "input channels must be a positive multiple of 16 (got %d)",
                  cin);
  CVQPG_CHECK_ARG(cout > 0 && cout % 16 == 0,
                  "output channels must be a positive multiple of 16 (got %d)", cout);
  CVQPG_CHECK_ARG(ksize == 1 || ksize == 3, "kernel size must be 1 or 3 (got %d)", ksize);
  return 0;
}

}  // namespace cvqpg

using namespace cvqpg;

extern "C" int cvqpg_nn_conv(const uint16_t* x, int batch, int height, int width, int cin,
                             const uint16_t* w, const float* bias, int cout, int ksize, int flags,
                             const uint16_t* addend, uint16_t* y_bf16, uint16_t* y2_bf16,
                             float* y_f32, void* stream) {
  if (int rc = check_conv(batch, height, width, cin, cout, ksize)) return rc;
  CVQPG_CHECK_ARG(x && w, "NULL input/filter");
  CVQPG_CHECK_ARG(y_bf16 || y_f32, "no output");
  CVQPG_CHECK_ARG(!y2_bf16 || (addend && y_bf16), "the second output needs addend and y_bf16");
  CVQPG_CHECK_ARG((flags & ~(CONV_RELU | CONV_ACCUMULATE)) == 0, "unknown flags %d", flags);
  const int M = batch * height * width;
  dim3 grid(ceil_div(M, 64), ceil_div(cout, 64));
  hipStream_t s = (hipStream_t)stream;
  if (ksize == 3)
    hipLaunchKernelGGL(k_conv<3>, grid, dim3(256), 0, s, x, batch, height, width, cin, w, bias,
                       cout, flags, addend, y_bf16, y2_bf16, y_f32);
  else
    hipLaunchKernelGGL(k_conv<1>, grid, dim3(256), 0, s, x, batch, height, width, cin, w, bias,
                       cout, flags, addend, y_bf16, y2_bf16, y_f32);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

static int wgrad_chunks(int batch, int height, int width) {
  return ceil_div((int64_t)batch * height * width, WG_PIX);
}

extern "C" size_t cvqpg_nn_wgrad_workspace_size(int batch, int height, int width, int cin,
                                               int cout, int ksize) {
  const size_t chunks = (size_t)wgrad_chunks(batch, height, width);
  return chunks * ((size_t)cout * ksize * ksize * cin + 4 * (size_t)cout) * sizeof(float) + 256;
}

extern "C" int cvqpg_nn_conv_wgrad(const uint16_t* x, int batch, int height, int width, int cin,
                                   const uint16_t* dy, int cout, int ksize, float* dw, float* db,
                                   void* ws, size_t ws_bytes, void* stream) {
  if (int rc = check_conv(batch, height, width, cin, cout, ksize)) return rc;
  CVQPG_CHECK_ARG(x && dy && dw && ws, "NULL input/gradient/output/workspace");
  CVQPG_CHECK_ARG(ws_bytes >= cvqpg_nn_wgrad_workspace_size(batch, height, width, cin, cout, ksize),
                  "workspace too small");
  const int chunks = wgrad_chunks(batch, height, width);
  CVQPG_CHECK_ARG(chunks <= 65535, "too many pixels for one weight-gradient call");
  const int64_t nw = (int64_t)cout * ksize * ksize * cin;
  float* P = static_cast<float*>(ws);
  float* Q = P + (size_t)chunks * nw;
  const int blocks = (cout / 16) * (ksize * ksize * cin / 16);
  hipStream_t s = (hipStream_t)stream;
  dim3 grid(ceil_div(blocks, 4), chunks);
  if (ksize == 3)
    hipLaunchKernelGGL(k_conv_wgrad<3>, grid, dim3(256), 0, s, x, batch, height, width, cin, dy,
                       cout, P, Q);
  else
    hipLaunchKernelGGL(k_conv_wgrad<1>, grid, dim3(256), 0, s, x, batch, height, width, cin, dy,
                       cout, P, Q);
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks_for(nw > cout ? nw : cout)), dim3(256), 0, s, P,
                     Q, chunks, nw, cout, dw, db);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_nn_flip_filter(const uint16_t* w, int cout, int cin, int ksize,
                                    uint16_t* w_flipped, void* stream) {
  CVQPG_CHECK_ARG(w && w_flipped && cout > 0 && cin > 0 && (ksize == 1 || ksize == 3),
                  "invalid filter arguments");
  const int64_t n = (int64_t)cout * ksize * ksize * cin;
  hipLaunchKernelGGL(k_flip, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, w, cout, cin,
                     ksize * ksize, w_flipped);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_nn_relu_backward(const float* dpost, const uint16_t* post, int64_t n,
                                      uint16_t* dpre, void* stream) {
  CVQPG_CHECK_ARG(dpost && dpre && n >= 0, "invalid arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, dpost,
                     post, n, dpre);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_nn_pool2(const uint16_t* x, int batch, int height, int width, int channels,
                              uint16_t* y, void* stream) {
  CVQPG_CHECK_ARG(x && y && batch > 0 && height > 0 && width > 0 && channels > 0,
                  "invalid pooling arguments");
  const int64_t n = (int64_t)batch * ((height + 1) / 2) * ((width + 1) / 2) * channels;
  hipLaunchKernelGGL(k_pool2, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, x, batch,
                     height, width, channels, y);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_nn_pool2_backward(const float* dy, int batch, int height, int width,
                                       int channels, float* dx, int accumulate, void* stream) {
  CVQPG_CHECK_ARG(dy && dx && batch > 0 && height > 0 && width > 0 && channels > 0,
                  "invalid pooling arguments");
  const int64_t n = (int64_t)batch * height * width * channels;
  hipLaunchKernelGGL(k_pool2_bwd, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, dy,
                     batch, height, width, channels, dx, accumulate);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_nn_upsample2(const uint16_t* x, int batch, int src_height, int src_width,
                                  int channels, int height, int width, uint16_t* y, void* stream) {
  CVQPG_CHECK_ARG(x && y && batch > 0 && channels > 0 && src_height == (height + 1) / 2 &&
                      src_width == (width + 1) / 2 && height > 0 && width > 0,
                  "invalid upsampling arguments (source must be ceil(target / 2))");
  const int64_t n = (int64_t)batch * height * width * channels;
  hipLaunchKernelGGL(k_up2, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, x, batch,
                     src_height, src_width, channels, height, width, y);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_nn_upsample2_backward(const float* dy, int batch, int src_height,
                                           int src_width, int channels, int height, int width,
                                           float* dx, int accumulate, void* stream) {
  CVQPG_CHECK_ARG(dy && dx && batch > 0 && channels > 0 && src_height == (height + 1) / 2 &&
                      src_width == (width + 1) / 2 && height > 0 && width > 0,
                  "invalid upsampling arguments (source must be ceil(target / 2))");
  const int64_t n = (int64_t)batch * src_height * src_width * channels;
  hipLaunchKernelGGL(k_up2_bwd, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, dy,
                     batch, src_height, src_width, channels, height, width, dx, accumulate);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_nn_tiles_from_image(const float* img, int batch, int channels, int height,
                                         int width, uint16_t* tiles, void* stream) {
  CVQPG_CHECK_ARG(img && tiles && batch > 0 && channels > 0 && height > 0 && width > 0,
                  "invalid image arguments");
  const int Ty = ceil_div(height, 16), Tx = ceil_div(width, 16);
  const int64_t n = (int64_t)batch * Ty * Tx * channels * 256;
  hipLaunchKernelGGL(k_tiles_from_image, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream,
                     img, batch, channels, height, width, Ty, Tx, tiles);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" size_t cvqpg_nn_wmse_workspace_size(int64_t n) {
  return (size_t)blocks_for(n) * sizeof(float);
}

extern "C" int cvqpg_nn_wmse(const float* out, const float* target, const float* weight, int64_t n,
                             float scale, float* dout, void* ws, size_t ws_bytes, float* loss,
                             void* stream) {
  CVQPG_CHECK_ARG(out && target && dout && ws && loss && n > 0, "invalid loss arguments");
  CVQPG_CHECK_ARG(ws_bytes >= cvqpg_nn_wmse_workspace_size(n), "workspace too small");
  hipStream_t s = (hipStream_t)stream;
  const int nb = blocks_for(n);
  hipLaunchKernelGGL(k_wmse, dim3(nb), dim3(256), 0, s, out, target, weight, n, scale, dout,
                     (float*)ws);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, (const float*)ws, nb, loss);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_nn_adam(float* w, uint16_t* w_bf16, const float* grad, float* m, float* v,
                             int64_t n, float lr, float beta1, float beta2, float eps, int step,
                             void* stream) {
  CVQPG_CHECK_ARG(w && grad && m && v && n >= 0 && step >= 1, "invalid Adam arguments");
  if (n == 0) return 0;
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, w, w_bf16,
                     grad, m, v, n, lr, beta1, beta2, eps, c1, c2);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_nn_to_bf16(const float* src, int64_t n, uint16_t* dst, void* stream) {
  CVQPG_CHECK_ARG(src && dst && n >= 0, "invalid arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, src, n,
                     dst);
  CVQPG_LAUNCH_CHECK();
  return 0;
}
