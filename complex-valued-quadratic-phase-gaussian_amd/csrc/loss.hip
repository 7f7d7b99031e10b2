// Multi-plane defocus loss (three-term masked L2 + SSIM) and its gradient for gfx950.
//
// Reference behaviour:
//   per-plane loss assembly        train_2d_gaussian_rgb.py:191-197
//   multiplane_loss_odak.__call__  utils/propagator.py:417-449 (weights [1, 1, 1],
//                                  data_utils.py:75; target noise
//                                  randn * 1e-3 * (max T - min T) when inject_noise)
//   GaussianLoss (SSIM term)       utils/data_utils.py:22-41: 0.025 * 0.2 * (1 - SSIM),
//                                  SSIM = 11-tap sigma-1.5 Gaussian window, K = (0.01,
//                                  0.03), data range 1, valid convolution, mean over the
//                                  map and the channels (pytorch-msssim's algorithm)
//
// For plane d, channel c, pixel p with I = clamp(intensity, 0, 1), T the target,
// M the mask, T~ = T + noise (n = C H W pixels per plane):
//   L_d = 1/n sum (I - T)^2 + 1/n sum M (I - T~)^2 + 1/n sum T~^2 (I - T~)^2
//         + 0.005 (1 - SSIM_d)
//   dL/dI = 2/n [(I - T) + (M + T~^2)(I - T~)] + dSSIM-term/dI, times [0 <= raw <= 1]
// The SSIM adjoint is the transposed (full) Gaussian filter of three coefficient maps
//   A = dS/dmu_x (total), B = dS/dE[x^2], Cc = dS/dE[xy]
// so dSSIM/dI = G^T A + 2 I G^T B + T G^T Cc.
//
// One streaming kernel (k_loss_stream, described below) computes the SSIM maps, the
// coefficient maps, the transposed filters and the L2 terms in a single pass over the
// intensities, target and mask; block partial sums give the loss value without a host
// read-back (k_loss_total).
#include <mutex>

#include "common.h"
#include "cvqpg.h"

namespace cvqpg {

constexpr int WIN = 11;
constexpr float SSIM_C1 = 0.0001f;     // (0.01 * 1)^2
constexpr float SSIM_C2 = 0.0009f;     // (0.03 * 1)^2
constexpr float LAMBDA = 0.025f * 0.2f;

struct LossParams {
  int C, D, H, W, Hv, Wv;
  int own, rs;         // pixel columns per strip, pixel rows per segment (balanced)
  float win[WIN];
  float inv_n;         // 1 / (C H W)
  float ssim_coef;     // -LAMBDA / (C Hv Wv): dL/dS for every valid-map entry
  float noise_scale[16];  // 1e-3 * (max T_d - min T_d) per plane (0: no noise)
  unsigned long long seed;
};

// ---------------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) counter-based generator + Box-Muller
// ---------------------------------------------------------------------------------
__device__ __forceinline__ void philox_round(uint32_t& c0, uint32_t& c1, uint32_t& c2,
                                             uint32_t& c3, uint32_t k0, uint32_t k1) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
  c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
}

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

// One tap of an 11-tap filter chain: the chain starts with a product (no zeroed
// accumulator to set up) and continues with fused multiply-adds.  Every chain of the loss
// and of the target statistics is formed this way, so the two agree bit for bit.
__device__ __forceinline__ float tap(int b, float w, float x, float acc) {
  return b == 0 ? w * x : fmaf(w, x, acc);
}

// Raw buffer view of one plane (bytes): a load at an offset at or past the end returns 0,
// so a guarded load is one load from a selected offset, with no branch, no zeroed
// register and no 64-bit address arithmetic.  Word 3 is the gfx9-family raw-buffer format.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_view(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           0x00020000);
}
constexpr uint32_t OUTSIDE = 0x80000000u;   // an offset past every plane
__device__ __forceinline__ float load_f32(__amdgpu_buffer_rsrc_t v, uint32_t off) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(v, (int)off, 0, 0));
}
__device__ __forceinline__ float2 load_f32x2(__amdgpu_buffer_rsrc_t v, uint32_t off) {
  const auto u = __builtin_amdgcn_raw_buffer_load_b64(v, (int)off, 0, 0);
  return make_float2(__uint_as_float(u[0]), __uint_as_float(u[1]));
}

// target noise: one Philox4x32-10 draw gives the four normals of pixels (y, x) with y in
// 4 g .. 4 g + 3 (two Box-Muller pairs); the counter is (g W + x, plane, step).
__device__ __forceinline__ void normals4(unsigned long long seed, uint32_t step, uint32_t plane,
                                         uint32_t idx, float (&z)[4]) {
  uint32_t c0 = idx, c1 = plane, c2 = step, c3 = 0x5EED0001u;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  // the key schedule is formed here, on the scalar unit, at every draw: hoisted out of the
  // streaming loop its twenty round keys did not fit the scalar registers and were kept in
  // vector-register lanes, read back one by one
  asm volatile("" : "+s"(k0), "+s"(k1));
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    philox_round(c0, c1, c2, c3, k0, k1);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  const float u1 = ((float)(c0 >> 8) + 0.5f) * (1.0f / 16777216.0f);
  const float u2 = (float)(c1 >> 8) * (1.0f / 16777216.0f);
  const float u3 = ((float)(c2 >> 8) + 0.5f) * (1.0f / 16777216.0f);
  const float u4 = (float)(c3 >> 8) * (1.0f / 16777216.0f);
  // Box-Muller on the hardware units: -2 ln u = -2 ln2 log2 u (v_log_f32), v_sqrt_f32,
  // and sin / cos of 2 pi u from v_sin_f32 / v_cos_f32, which take revolutions
  constexpr float M2LN2 = -1.3862943611198906f;
  const float r1 = __builtin_amdgcn_sqrtf(M2LN2 * __builtin_amdgcn_logf(u1));
  const float r2 = __builtin_amdgcn_sqrtf(M2LN2 * __builtin_amdgcn_logf(u3));
  z[0] = r1 * __builtin_amdgcn_cosf(u2);
  z[1] = r1 * __builtin_amdgcn_sinf(u2);
  z[2] = r2 * __builtin_amdgcn_cosf(u4);
  z[3] = r2 * __builtin_amdgcn_sinf(u4);
}

// ---------------------------------------------------------------------------------
// One streaming pass per (plane, channel, column strip, row segment).
//
// Thread i of a workgroup of NTH owns SSIM-map column mc = x0 - 10 + i and, for i >= 10,
// pixel column mc.  The workgroup walks down the rows of its segment; at each step it
//   1. stages input row y (clamped intensity and target, NTH + 10 columns) in LDS; the
//      row was loaded into registers during the previous step,
//   2. filters it horizontally into the five moments and pushes them into an 11-row
//      register ring,
//   3. filters the ring vertically for map row q = y - 10 -> SSIM and the coefficient
//      maps A, B, Cc of that row (never stored to HBM: staged in LDS for one step),
//   4. applies the transposed horizontal filter to the map row and pushes the result
//      into a second 11-row register ring (private to the thread's column),
//   5. for pixel row p = q, applies the transposed vertical filter from that ring and
//      adds the L2 terms, the clamp mask and the target noise: dL/dI.  The pixel's
//      intensity, target and mask were loaded during the previous step.
// The rings are indexed by the step modulo 11, which the 11-fold unrolled step loop
// makes a compile-time choice.  Pixel rows [p0, p0 + rs) need map rows [p0 - 10,
// p0 + rs) and input rows [p0 - 10, p0 + rs + 10): rs + 20 steps.  The host splits the
// columns into equal strips of own <= OWN columns (threads i >= own + 10 of a narrower
// strip only stage input columns) and the rows into equal segments, choosing the
// segment count so that the workgroups fill whole passes of the GPU (loss_split).
// ---------------------------------------------------------------------------------
#ifndef CVQPG_LOSS_NTH
#define CVQPG_LOSS_NTH 256
#endif
constexpr int NTH = CVQPG_LOSS_NTH;   // threads = map columns per strip
constexpr int OWN = NTH - (WIN - 1);  // most pixel columns per strip
#ifndef CVQPG_LOSS_RS
#define CVQPG_LOSS_RS 192
#endif
constexpr int RS = CVQPG_LOSS_RS;     // segment rows when the device cannot be queried
#ifndef CVQPG_LOSS_RS_MIN
#define CVQPG_LOSS_RS_MIN 100
#endif
constexpr int RS_MIN = CVQPG_LOSS_RS_MIN;  // segments of at least ~RS_MIN rows: at most
                                           // 20 / RS_MIN of the steps repeat halo rows
static_assert(RS % 4 == 0, "segments hold whole groups of four rows");
constexpr int NIN = NTH + WIN - 1;    // input columns staged per row

// waves per SIMD the register allocation must leave room for (1: no constraint; the
// kernel reaches 4 with precomputed target statistics and 3 without, spill-free; 5 spills)
#ifndef CVQPG_LOSS_WPE
#define CVQPG_LOSS_WPE 1
#endif
// TS: the target's window statistics (mu_t, E[t^2]) of every map entry come precomputed
// (k_target_stats, bit-identical to filtering them here); only the three moments of the
// intensity (x, x^2, x t) are then filtered per step.
// L2 = false leaves the three L2 terms (and the mask and noise) out: the SSIM term alone,
// GaussianLoss of utils/data_utils.py:22-41.
template <bool TS, bool L2 = true>
__global__ __launch_bounds__(NTH, CVQPG_LOSS_WPE) void k_loss_stream(
    const float* __restrict__ inten, const float* __restrict__ target,
    const float* __restrict__ mask, const float* __restrict__ noise,
    const float2* __restrict__ tstats, const int* __restrict__ step_counter, LossParams P,
    float* __restrict__ grad, float* __restrict__ ssim_part, float* __restrict__ mse_part) {
  constexpr int NM = TS ? 3 : 5;     // filtered moments: x, x^2, x t (, t, t^2)
  __shared__ float smo[2][NM][NIN];  // staged input row
  __shared__ float smap[2][3][NTH];
  __shared__ float red[2][NTH];
  // raw (intensity, target) of the thread's input column mc for the last 11 input rows,
  // by step modulo 11: pixel row y - 10 of a step was staged as input row 10 steps
  // earlier by the same thread, so its values are not loaded from HBM a second time
  __shared__ float2 pring[WIN][NTH];
  const int i = threadIdx.x;
  const int x0 = blockIdx.x * P.own;
  const int p0 = blockIdx.y * P.rs;
  const int dc = blockIdx.z;
  const int d = dc / P.C, c = dc % P.C;
  const int mc = x0 - (WIN - 1) + i;              // map column (and pixel column, i >= 10)
  const bool own_col = i >= WIN - 1 && i < P.own + WIN - 1;
  const bool map_col = mc >= 0 && mc < P.Wv;
  const bool pix_col = own_col && mc < P.W;
  // plane offsets in 64 bits, offsets inside a plane (< 2^31 elements) in 32 bits
  const float* inp = inten + (size_t)dc * P.H * P.W;
  const float* tgp = target + (size_t)dc * P.H * P.W;
  const float* mkp = mask + (size_t)dc * P.H * P.W;
  const float* nzp = noise ? noise + (size_t)dc * P.H * P.W : nullptr;
  float* grp = grad + (size_t)dc * P.H * P.W;
  const int Wd = P.W;
  const uint32_t step = step_counter ? (uint32_t)step_counter[0] : 0u;
  const float ns = P.noise_scale[d];
  const int p_end = min(p0 + P.rs, P.H);
  const int ys = p0 - (WIN - 1);                  // first input row
  const int nsteps = P.rs + 2 * (WIN - 1);
  const uint32_t hq = (uint32_t)((P.H + 3) / 4);
  // the input columns this thread stages: x0 - 10 + i and, for i < 10, x0 - 10 + i + NTH
  const int gx0 = mc, gx1 = mc + NTH;
  const bool two = i < WIN - 1;
  float pa0 = 0.f, pb0 = 0.f, pa1 = 0.f, pb1 = 0.f;   // next input row (raw inten, target)
  float qM = 0.f, qN = 0.f;                           // next pixel row values
  // out-of-range rows and columns read 0 (buffer loads past the plane's end)
  const uint32_t pbytes = (uint32_t)(P.H * P.W) * 4u;
  const __amdgpu_buffer_rsrc_t vin = plane_view(inp, pbytes), vtg = plane_view(tgp, pbytes);
  const __amdgpu_buffer_rsrc_t vmk = plane_view(mkp, pbytes);
  const __amdgpu_buffer_rsrc_t vnz = plane_view(nzp ? nzp : inp, nzp ? pbytes : 0u);
  auto load_row = [&](int y) {
    const bool yin = y >= 0 && y < P.H;
    const uint32_t r = (uint32_t)(y * Wd);
    const uint32_t o0 = yin && gx0 >= 0 && gx0 < Wd ? (r + (uint32_t)gx0) * 4u : OUTSIDE;
    pa0 = load_f32(vin, o0);
    pb0 = load_f32(vtg, o0);
    if (two) {
      const uint32_t o1 = yin && gx1 < Wd ? (r + (uint32_t)gx1) * 4u : OUTSIDE;
      pa1 = load_f32(vin, o1);
      pb1 = load_f32(vtg, o1);
    }
  };
  auto load_pix = [&](int pr) {
    const uint32_t o = pix_col && pr >= p0 && pr < p_end ? (uint32_t)(pr * Wd + mc) * 4u : OUTSIDE;
    qM = load_f32(vmk, o);
    if (nzp) qN = load_f32(vnz, o);
  };
  // target statistics of map row q at this thread's map column (TS)
  const float2* tsp = TS ? tstats + (size_t)dc * P.Hv * P.Wv : nullptr;
  const __amdgpu_buffer_rsrc_t vst =
      plane_view(TS ? (const void*)tsp : (const void*)inp, TS ? (uint32_t)(P.Hv * P.Wv) * 8u : 0u);
  float2 sN = make_float2(0.f, 0.f);
  auto load_stats = [&](int q) {
    if (TS) {
      const uint32_t o = q >= 0 && q < P.Hv && map_col ? (uint32_t)(q * P.Wv + mc) * 8u : OUTSIDE;
      sN = load_f32x2(vst, o);
    }
  };
  float hm[WIN][NM];   // moment ring: row-filtered moments of input rows
#pragma unroll
  for (int k = 0; k < WIN; ++k)
#pragma unroll
    for (int m = 0; m < NM; ++m) hm[k][m] = 0.f;
  float gring[WIN][3];  // row-filtered coefficient-map ring of the thread's column
#pragma unroll
  for (int k = 0; k < WIN; ++k)
#pragma unroll
    for (int m = 0; m < 3; ++m) gring[k][m] = 0.f;
  float z4[4] = {0.f, 0.f, 0.f, 0.f};
  float ssum = 0.f, msum = 0.f;
  // The gradient of a pixel row is stored at the start of the next step, before that step
  // issues its loads: the loads of a step are waited for at the next step, and a store
  // issued after them would be waited for with them (the memory counter is in order), so
  // every step would wait out a store's round trip.
  float gpend = 0.f;
  int ipend = -1;
  load_row(ys);
  load_stats(ys - (WIN - 1));

  for (int base = 0; base < nsteps; base += WIN) {
#pragma unroll
    for (int k = 0; k < WIN; ++k) {
      const int sidx = base + k;
      if (sidx >= nsteps) break;
      const int y = ys + sidx;
      const int buf = sidx & 1;
      // 1. stage input row y (loaded last step), then start loading row y + 1 and the
      //    pixel values of pixel row y - 9 (used next step)
      {
        const float a = clamp01(pa0), b = pb0;
        smo[buf][0][i] = a; smo[buf][1][i] = a * a; smo[buf][2][i] = a * b;
        pring[k][i] = make_float2(pa0, pb0);
        if (!TS) { smo[buf][NM - 2][i] = b; smo[buf][NM - 1][i] = b * b; }
        if (two) {
          const float a1 = clamp01(pa1), b1 = pb1;
          smo[buf][0][i + NTH] = a1; smo[buf][1][i + NTH] = a1 * a1;
          smo[buf][2][i + NTH] = a1 * b1;
          if (!TS) { smo[buf][NM - 2][i + NTH] = b1; smo[buf][NM - 1][i + NTH] = b1 * b1; }
        }
      }
      const float cM = qM, cN = qN;
      const float2 cS = sN;
      if (ipend >= 0) grp[ipend] = gpend;
      ipend = -1;
      load_row(y + 1);
      load_pix(y - (WIN - 1) + 1);
      load_stats(y - (WIN - 1) + 1);
      __syncthreads();
      // 2. horizontal moments of map column mc (input columns mc .. mc + 10)
      {
        float m[NM];
#pragma unroll
        for (int b = 0; b < WIN; ++b) {
          const float w = P.win[b];
#pragma unroll
          for (int j = 0; j < NM; ++j) m[j] = tap(b, w, smo[buf][j][i + b], m[j]);
        }
#pragma unroll
        for (int j = 0; j < NM; ++j) hm[k][j] = m[j];
      }
      if (sidx < WIN - 1) continue;          // ring not yet full: no map row
      // 3. map row q = y - 10 from input rows q .. q + 10 (ring slots k + 1 + a mod 11)
      const int q = y - (WIN - 1);
      float cA = 0.f, cB = 0.f, cC = 0.f;
      if (q >= 0 && q < P.Hv && map_col) {
        float mu1 = 0.f, mu2 = 0.f, e11 = 0.f, e22 = 0.f, e12 = 0.f;
#pragma unroll
        for (int a = 0; a < WIN; ++a) {
          const int sl = (k + 1 + a) % WIN;
          const float w = P.win[a];
          mu1 = tap(a, w, hm[sl][0], mu1);
          e11 = tap(a, w, hm[sl][1], e11);
          e12 = tap(a, w, hm[sl][2], e12);
          if (!TS) {
            mu2 = tap(a, w, hm[sl][NM - 2], mu2);
            e22 = tap(a, w, hm[sl][NM - 1], e22);
          }
        }
        if (TS) {
          mu2 = cS.x;
          e22 = cS.y;
        }
        const float s11 = e11 - mu1 * mu1, s22 = e22 - mu2 * mu2, s12 = e12 - mu1 * mu2;
        const float A1 = 2.f * mu1 * mu2 + SSIM_C1, A2 = 2.f * s12 + SSIM_C2;
        const float B1 = mu1 * mu1 + mu2 * mu2 + SSIM_C1, B2 = s11 + s22 + SSIM_C2;
        // one hardware reciprocal (1 ulp) of B1 B2 > 0; S / A1 and S / A2 are written
        // as A2 / (B1 B2) and A1 / (B1 B2), so A1 or A2 = 0 needs no special case
        const float rB = __builtin_amdgcn_rcpf(B1 * B2);
        const float S = (A1 * A2) * rB;
        if (own_col && q >= p0) ssum += S;
        const float kc = P.ssim_coef;
        // total derivative with respect to mu_x, E[x^2], E[xy]:
        //   dS/dmu_x = 2 mu_y (A2 - A1) / (B1 B2) + 2 mu_x S (B1 - B2) / (B1 B2)
        //   dS/dE[x^2] = -S / B2,  dS/dE[xy] = 2 A1 / (B1 B2)
        cA = 2.f * kc * rB * (mu2 * (A2 - A1) + mu1 * S * (B1 - B2));
        cB = -kc * S * B1 * rB;
        cC = 2.f * kc * A1 * rB;
      }
      smap[buf][0][i] = cA;
      smap[buf][1][i] = cB;
      smap[buf][2][i] = cC;
      __syncthreads();
      // 4. transposed horizontal filter: map columns mc - 10 .. mc (threads i - 10 .. i)
      if (!own_col) continue;
      {
        float g0 = 0.f, g1 = 0.f, g2 = 0.f;
#pragma unroll
        for (int b = 0; b < WIN; ++b) {
          const float w = P.win[b];
          g0 = tap(b, w, smap[buf][0][i - b], g0);
          g1 = tap(b, w, smap[buf][1][i - b], g1);
          g2 = tap(b, w, smap[buf][2][i - b], g2);
        }
        gring[k][0] = g0; gring[k][1] = g1; gring[k][2] = g2;
      }
      // 5. pixel row p = q: transposed vertical filter over map rows p - a (slots k - a)
      const int pr = q;
      if (pr < p0 || pr >= p_end) continue;
      if (L2 && ns != 0.f && !noise && ((pr - p0) & 3) == 0 && pix_col)
        normals4(P.seed, step, (uint32_t)d,
                 ((uint32_t)c * hq + (uint32_t)(pr >> 2)) * (uint32_t)Wd + (uint32_t)mc, z4);
      if (!pix_col) continue;
      const float2 pit = pring[(k + 1) % WIN][i];   // staged at step sidx - 10: row pr
      const float cI = pit.x, cT = pit.y;
      float gA = 0.f, gB = 0.f, gC = 0.f;
#pragma unroll
      for (int a = 0; a < WIN; ++a) {
        const int sl = (k - a + WIN) % WIN;
        const float w = P.win[a];
        gA = tap(a, w, gring[sl][0], gA);
        gB = tap(a, w, gring[sl][1], gB);
        gC = tap(a, w, gring[sl][2], gC);
      }
      const float I = clamp01(cI);
      float gM = 0.f;
      if constexpr (L2) {
        float nz = 0.f;
        if (ns != 0.f) {
          const int l = (pr - p0) & 3;
          nz = noise ? cN : (l == 0 ? z4[0] : l == 1 ? z4[1] : l == 2 ? z4[2] : z4[3]);
          nz *= ns;
        }
        const float Tn = cT + nz;
        const float e1 = I - cT, e2 = I - Tn;
        const float wgt = cM * cM + Tn * Tn;
        msum += e1 * e1 + wgt * e2 * e2;
        gM = 2.f * P.inv_n * (e1 + wgt * e2);
      }
      const float gS = gA + 2.f * I * gB + cT * gC;
      gpend = (cI >= 0.f && cI <= 1.f) ? gM + gS : 0.f;
      ipend = pr * Wd + mc;
    }
  }
  if (ipend >= 0) grp[ipend] = gpend;
  red[0][i] = ssum;
  red[1][i] = msum;
  __syncthreads();
  for (int off = NTH / 2; off > 0; off >>= 1) {
    if (i < off) {
      red[0][i] += red[0][i + off];
      red[1][i] += red[1][i + off];
    }
    __syncthreads();
  }
  if (i == 0) {
    const size_t b = ((size_t)dc * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    ssim_part[b] = red[0][0];
    mse_part[b] = red[1][0];
  }
}

// Window statistics of the target for k_loss_stream<true>: (mu_t, E[t^2]) of every valid
// map entry [D C][Hv][Wv], with the streaming kernel's filter chains (tap; row filter over
// columns mc .. mc + 10, then column filter over rows q .. q + 10), so that the loss and
// its gradient are bit-identical to filtering the target in the streaming kernel.
// grid (ceil(Wv / 256), Hv, D C)
__global__ __launch_bounds__(256) void k_target_stats(const float* __restrict__ target,
                                                      LossParams P, float2* __restrict__ stats) {
  const int mc = blockIdx.x * 256 + threadIdx.x, q = blockIdx.y, dc = blockIdx.z;
  if (mc >= P.Wv) return;
  const float* tp = target + (size_t)dc * P.H * P.W;
  float mu = 0.f, e = 0.f;
#pragma unroll
  for (int a = 0; a < WIN; ++a) {
    const float* row = tp + (size_t)(q + a) * P.W + mc;
    float h1 = 0.f, h2 = 0.f;
#pragma unroll
    for (int b = 0; b < WIN; ++b) {
      const float t = row[b];
      h1 = tap(b, P.win[b], t, h1);
      h2 = tap(b, P.win[b], t * t, h2);
    }
    mu = tap(a, P.win[a], h1, mu);
    e = tap(a, P.win[a], h2, e);
  }
  stats[((size_t)dc * P.Hv + q) * P.Wv + mc] = make_float2(mu, e);
}

// The target noise the streaming kernel draws in-kernel, as a tensor [D][C][H][W] of unit
// normals: thread (x, row group g, plane-channel) makes the same Philox draw for pixels
// (4 g .. 4 g + 3, x), so passing the tensor as `noise` gives the same loss and gradient
// bit for bit.  grid (ceil(W / 256), ceil(H / 4), D C)
__global__ __launch_bounds__(256) void k_loss_noise(unsigned long long seed,
                                                    const int* __restrict__ step_counter,
                                                    int C, int H, int W,
                                                    float* __restrict__ noise) {
  const int x = blockIdx.x * 256 + threadIdx.x, g = blockIdx.y, dc = blockIdx.z;
  if (x >= W) return;
  const int d = dc / C, c = dc % C;
  const uint32_t step = step_counter ? (uint32_t)step_counter[0] : 0u;
  const uint32_t hq = (uint32_t)((H + 3) / 4);
  float z4[4];
  normals4(seed, step, (uint32_t)d, ((uint32_t)c * hq + (uint32_t)g) * (uint32_t)W + (uint32_t)x,
           z4);
  float* np = noise + (size_t)dc * H * W;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const int y = 4 * g + l;
    if (y < H) np[(size_t)y * W + x] = z4[l];
  }
}

// loss = sum_d [ mse_d / n + LAMBDA (1 - SSIM_d) ]; one workgroup, fixed summation order
__global__ __launch_bounds__(1024) void k_loss_total(const float* __restrict__ ssim_part,
                                                     int n_ssim, const float* __restrict__ mse_part,
                                                     int n_mse, LossParams P,
                                                     float* __restrict__ loss) {
  __shared__ double red[1024];
  const int t = threadIdx.x;
  double s[4] = {0.0, 0.0, 0.0, 0.0}, m[4] = {0.0, 0.0, 0.0, 0.0};
  int i = t;
  for (; i + 3 * 1024 < n_ssim; i += 4 * 1024) {
#pragma unroll
    for (int u = 0; u < 4; ++u) s[u] += ssim_part[i + u * 1024];
  }
  for (; i < n_ssim; i += 1024) s[0] += ssim_part[i];
  i = t;
  for (; i + 3 * 1024 < n_mse; i += 4 * 1024) {
#pragma unroll
    for (int u = 0; u < 4; ++u) m[u] += mse_part[i + u * 1024];
  }
  for (; i < n_mse; i += 1024) m[0] += mse_part[i];
  red[t] = (double)P.inv_n * ((m[0] + m[1]) + (m[2] + m[3])) -
           (double)LAMBDA / ((double)P.C * P.Hv * P.Wv) * ((s[0] + s[1]) + (s[2] + s[3]));
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  if (t == 0) loss[0] = (float)(red[0] + (double)LAMBDA * P.D);
}

struct LossWS {
  float* ssim_part;
  float* mse_part;
  size_t bytes;
};

// Workgroups of k_loss_stream<TS> the current device runs at once (CUs x workgroups per
// CU at the kernel's register and LDS use), queried once per device; 0 if there is no
// device.
template <bool TS>
static int loss_slots() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  if (!cached[dev]) {
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_loss_stream<TS>, NTH, 0) !=
            hipSuccess)
      return 0;
    cached[dev] = cus * per;
  }
  return cached[dev];
}

static int round4(int v) { return (v + 3) & ~3; }

// Strips of own <= OWN columns and segments of rs rows (a multiple of 4, the rows that share
// one Philox draw), each as equal as the counts allow.  Every workgroup walks rs + 20 row
// steps, so the kernel takes about (passes of the GPU) x (rs + 20) step times: the segment
// count (at most ceil(H / RS_MIN)) minimising that product is chosen, the fewest segments
// on a tie.
// With 256 CUs holding 3 workgroups each (the kernel filtering the target itself), a
// 1920x1080 frame of 4 planes x 3 channels splits into 8 strips x 8 segments of 136 rows:
// 768 workgroups, one full pass; with 4 per CU (precomputed target statistics), into
// 8 strips x 10 segments of 108 rows: 960 workgroups, one pass.
static void loss_split(int slots, int DC, int H, int W, int& own, int& rs) {
  const int strips = ceil_div(W, OWN);
  own = ceil_div(W, strips);
  if (slots <= 0) {
    rs = round4(ceil_div(H, ceil_div(H, RS)));
    return;
  }
  long best = -1;
  for (int n = 1; n <= max(1, ceil_div(H, RS_MIN)); ++n) {
    const int r = round4(ceil_div(H, n));
    const long wgs = (long)strips * ceil_div(H, r) * DC;
    const long cost = (long)ceil_div(wgs, (long)slots) * (r + 2 * (WIN - 1));
    if (best < 0 || cost < best) {
      best = cost;
      rs = r;
    }
  }
}

// most workgroups any split launches (sizes the partial-sum workspace)
static int loss_max_blocks(int C, int D, int H, int W) {
  return ceil_div(W, OWN) * max(1, ceil_div(H, RS_MIN)) * D * C;
}

static LossWS carve_loss(void* base, int C, int D, int H, int W) {
  Carver cv(base);
  LossWS w;
  const size_t nb = (size_t)loss_max_blocks(C, D, H, W);
  w.ssim_part = cv.take<float>(nb);
  w.mse_part = cv.take<float>(nb);
  w.bytes = cv.bytes();
  return w;
}

}  // namespace cvqpg

using namespace cvqpg;

extern "C" size_t cvqpg_loss_workspace_size(int channels, int planes, int height, int width) {
  if (height < WIN || width < WIN) return 0;
  return carve_loss(nullptr, channels, planes, height, width).bytes;
}

static LossParams loss_params(int channels, int planes, int height, int width,
                              const float* noise_scale, uint64_t seed) {
  LossParams P;
  P.C = channels;
  P.D = planes;
  P.H = height;
  P.W = width;
  P.Hv = height - WIN + 1;
  P.Wv = width - WIN + 1;
  P.own = P.rs = 0;
  {
    double g[WIN], sum = 0.0;
    for (int i = 0; i < WIN; ++i) {
      const double c = i - WIN / 2;
      g[i] = exp(-(c * c) / (2.0 * 1.5 * 1.5));
      sum += g[i];
    }
    for (int i = 0; i < WIN; ++i) P.win[i] = (float)(g[i] / sum);
  }
  P.inv_n = (float)(1.0 / ((double)channels * height * width));
  P.ssim_coef = (float)(-(double)LAMBDA / ((double)channels * P.Hv * P.Wv));
  for (int d = 0; d < 16; ++d) P.noise_scale[d] = (noise_scale && d < planes) ? noise_scale[d] : 0.f;
  P.seed = seed;
  return P;
}

extern "C" int cvqpg_loss_noise(uint64_t seed, const int32_t* step_counter, int channels,
                                int planes, int height, int width, float* noise, void* stream) {
  CVQPG_CHECK_ARG(channels >= 1 && channels <= 3, "channels must be 1..3");
  CVQPG_CHECK_ARG(planes >= 1 && planes <= 16, "planes must be 1..16");
  CVQPG_CHECK_ARG(height >= 1 && width >= 1, "invalid image size %dx%d", width, height);
  CVQPG_CHECK_ARG(noise, "NULL argument pointer");
  hipLaunchKernelGGL(k_loss_noise, dim3(ceil_div(width, 256), ceil_div(height, 4),
                                        planes * channels),
                     dim3(256), 0, (hipStream_t)stream, (unsigned long long)seed, step_counter,
                     channels, height, width, noise);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" size_t cvqpg_loss_target_stats_size(int channels, int planes, int height,
                                               int width) {
  if (height < WIN || width < WIN || channels < 1 || planes < 1) return 0;
  return (size_t)planes * channels * (height - WIN + 1) * (width - WIN + 1) * sizeof(float2);
}

extern "C" int cvqpg_loss_target_stats(const float* target, int channels, int planes,
                                       int height, int width, float* stats, size_t stats_bytes,
                                       void* stream) {
  CVQPG_CHECK_ARG(channels >= 1 && channels <= 3, "channels must be 1..3");
  CVQPG_CHECK_ARG(planes >= 1 && planes <= 16, "planes must be 1..16");
  CVQPG_CHECK_ARG(height >= WIN && width >= WIN, "image smaller than the SSIM window");
  CVQPG_CHECK_ARG(target && stats, "NULL argument pointer");
  const size_t need = cvqpg_loss_target_stats_size(channels, planes, height, width);
  CVQPG_CHECK_ARG(stats_bytes >= need, "statistics buffer too small: %zu < %zu", stats_bytes, need);
  const LossParams P = loss_params(channels, planes, height, width, nullptr, 0);
  hipLaunchKernelGGL(k_target_stats, dim3(ceil_div(P.Wv, 256), P.Hv, planes * channels),
                     dim3(256), 0, (hipStream_t)stream, target, P,
                     reinterpret_cast<float2*>(stats));
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_loss_forward_backward_stats(
    const float* intensity, const float* target, const float* target_stats, const float* mask,
    const float* noise, const float* noise_scale, uint64_t seed, const int32_t* step_counter,
    int channels, int planes, int height, int width, void* ws, size_t ws_bytes,
    float* grad_intensity, float* loss, void* stream) {
  CVQPG_CHECK_ARG(channels >= 1 && channels <= 3, "channels must be 1..3");
  CVQPG_CHECK_ARG(planes >= 1 && planes <= 16, "planes must be 1..16");
  CVQPG_CHECK_ARG(height >= WIN && width >= WIN, "image smaller than the SSIM window");
  CVQPG_CHECK_ARG(intensity && target && mask && grad_intensity && loss && ws,
                  "NULL argument pointer");
  const LossWS need = carve_loss(nullptr, channels, planes, height, width);
  CVQPG_CHECK_ARG(ws_bytes >= need.bytes, "workspace too small: %zu < %zu", ws_bytes, need.bytes);
  hipStream_t s = (hipStream_t)stream;
  const LossWS w = carve_loss(ws, channels, planes, height, width);
  LossParams P = loss_params(channels, planes, height, width, noise_scale, seed);
  const float2* ts = reinterpret_cast<const float2*>(target_stats);
  loss_split(ts ? loss_slots<true>() : loss_slots<false>(), planes * channels, height, width,
             P.own, P.rs);
  const dim3 g(ceil_div(width, P.own), ceil_div(height, P.rs), planes * channels);
  const int nb = (int)(g.x * g.y * g.z);
  if (ts)
    hipLaunchKernelGGL(k_loss_stream<true>, g, dim3(NTH), 0, s, intensity, target, mask, noise,
                       ts, step_counter, P, grad_intensity, w.ssim_part, w.mse_part);
  else
    hipLaunchKernelGGL(k_loss_stream<false>, g, dim3(NTH), 0, s, intensity, target, mask,
                       noise, ts, step_counter, P, grad_intensity, w.ssim_part, w.mse_part);
  hipLaunchKernelGGL(k_loss_total, dim3(1), dim3(1024), 0, s, w.ssim_part, nb, w.mse_part, nb,
                     P, loss);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_loss_forward_backward(const float* intensity, const float* target,
                                           const float* mask, const float* noise,
                                           const float* noise_scale, uint64_t seed,
                                           const int32_t* step_counter, int channels,
                                           int planes, int height, int width, void* ws,
                                           size_t ws_bytes, float* grad_intensity, float* loss,
                                           void* stream) {
  return cvqpg_loss_forward_backward_stats(intensity, target, nullptr, mask, noise,
                                           noise_scale, seed, step_counter, channels, planes,
                                           height, width, ws, ws_bytes, grad_intensity, loss,
                                           stream);
}

extern "C" int cvqpg_ssim_term(const float* image, const float* target, int channels,
                               int height, int width, void* ws, size_t ws_bytes,
                               float* grad_image, float* loss, void* stream) {
  CVQPG_CHECK_ARG(channels >= 1 && channels <= 3, "channels must be 1..3");
  CVQPG_CHECK_ARG(height >= WIN && width >= WIN, "image smaller than the SSIM window");
  CVQPG_CHECK_ARG(image && target && grad_image && loss && ws, "NULL argument pointer");
  const LossWS need = carve_loss(nullptr, channels, 1, height, width);
  CVQPG_CHECK_ARG(ws_bytes >= need.bytes, "workspace too small: %zu < %zu", ws_bytes, need.bytes);
  hipStream_t s = (hipStream_t)stream;
  const LossWS w = carve_loss(ws, channels, 1, height, width);
  LossParams P = loss_params(channels, 1, height, width, nullptr, 0);
  loss_split(0, channels, height, width, P.own, P.rs);
  const dim3 g(ceil_div(width, P.own), ceil_div(height, P.rs), channels);
  const int nb = (int)(g.x * g.y * g.z);
  // the mask operand is never read without the L2 terms; the target stands in for it
  hipLaunchKernelGGL((k_loss_stream<false, false>), g, dim3(NTH), 0, s, image, target, target,
                     nullptr, nullptr, nullptr, P, grad_image, w.ssim_part, w.mse_part);
  hipLaunchKernelGGL(k_loss_total, dim3(1), dim3(1024), 0, s, w.ssim_part, nb, w.mse_part, nb,
                     P, loss);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

// ---------------------------------------------------------------------------------
// The same loss with its scratch (the per-workgroup partial sums) in a static device
// array instead of a caller's workspace, for callers that cannot allocate.  Every call on
// one GPU shares that GPU's array; each launch writes every partial it later reads, so
// nothing needs clearing between calls, but calls that overlap on one GPU race.
// ---------------------------------------------------------------------------------
namespace cvqpg {
constexpr size_t LOSS_SCRATCH_BYTES = 256 * 1024;
static __device__ __align__(256) unsigned char scratch[LOSS_SCRATCH_BYTES];

static std::mutex scratch_mu;
static void* scratch_ptr[64] = {nullptr};

// device address of this device's scratch (looked up once per device, with it current)
static int loss_scratch(void** out) {
  int dev = 0;
  CVQPG_HIP(hipGetDevice(&dev));
  CVQPG_CHECK_ARG(dev >= 0 && dev < 64, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    CVQPG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    scratch_ptr[dev] = p;
  }
  *out = scratch_ptr[dev];
  return 0;
}
}  // namespace cvqpg

extern "C" size_t cvqpg_loss_noalloc_max_bytes(void) { return LOSS_SCRATCH_BYTES; }

extern "C" int cvqpg_loss_forward_backward_noalloc(
    const float* intensity, const float* target, const float* target_stats, const float* mask,
    const float* noise, const float* noise_scale, uint64_t seed, const int32_t* step_counter,
    int channels, int planes, int height, int width, float* grad_intensity, float* loss,
    void* stream) {
  CVQPG_CHECK_ARG(channels >= 1 && channels <= 3, "channels must be 1..3");
  CVQPG_CHECK_ARG(planes >= 1 && planes <= 16, "planes must be 1..16");
  CVQPG_CHECK_ARG(height >= WIN && width >= WIN, "image smaller than the SSIM window");
  const size_t need = carve_loss(nullptr, channels, planes, height, width).bytes;
  CVQPG_CHECK_ARG(need <= LOSS_SCRATCH_BYTES,
                  "%d planes x %d channels of %dx%d need %zu bytes of loss scratch; the "
                  "no-allocation entry point holds %zu (use cvqpg_loss_forward_backward_stats "
                  "with a workspace)",
                  planes, channels, width, height, need, LOSS_SCRATCH_BYTES);
  void* ws = nullptr;
  if (int rc = loss_scratch(&ws)) return rc;
  return cvqpg_loss_forward_backward_stats(intensity, target, target_stats, mask, noise,
                                           noise_scale, seed, step_counter, channels, planes,
                                           height, width, ws, LOSS_SCRATCH_BYTES, grad_intensity,
                                           loss, stream);
}
