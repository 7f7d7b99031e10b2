// Band-limited angular spectrum propagation (forward and adjoint) for gfx950, fused with
// the guard-band padding, the amplitude/phase glue, the 2x zero padding, the centre crops
// and |.|^2.
//
// Reference behaviour:
//   soft_window_pad            model_2d_gaussian_rgb.py:186-241
//   reconstruct glue           train_2d_gaussian_rgb.py:89-95
//   propagator / custom()      utils/propagator.py:41-64, 248-290 (odak zero_pad,
//                              crop_center, BLASM transfer function)
//
// Per channel c the forward is three passes over HBM:
//   F1  rows   : build the padded, tapered, amplitude-clamped field on the fly from the
//                raster output, place it in the centre of a 2x row, FFT along the row;
//                only the pH non-zero rows of the 2x grid are transformed (X [pH, nv]).
//   F2  columns: load the pH non-zero entries of CG adjacent columns, FFT along the
//                column, keep the spectrum in registers, and for every plane d multiply
//                by H_d (phase evaluated on the fly from a host FP64 carrier and an FP32
//                residual, no cached transfer function), inverse FFT along the column and
//                store only the rows that survive the final crop (Y_d [oh, nv]).
//   F3  rows   : inverse FFT along the row, store only the cropped columns: the complex
//                field at the plane (kept for the adjoint) and its intensity.
// The adjoint runs the transposed passes in reverse (A3 rows, A2 columns with the
// conjugate transfer functions summed over planes, A1 rows + the glue adjoint), so the
// full 2x spectrum never goes to HBM in either direction.  The transform lengths of the
// standard workloads run on compile-time pipelines (k_*_fx, see fft::Pipe); any other
// size runs on the runtime-plan kernels.
#include <type_traits>

#include "common.h"
#include "cvqpg.h"
#include "fft.h"
#include "window.h"

namespace cvqpg {

constexpr int MAXD = 16;
#ifndef CVQPG_ROW_NT
#define CVQPG_ROW_NT 256
#endif
#ifndef CVQPG_COL_NT
#define CVQPG_COL_NT 512
#endif
// waves per SIMD the 2816-point column kernels are compiled for (measured best)
#ifndef CVQPG_WPE_F
#define CVQPG_WPE_F 4
#endif
#ifndef CVQPG_WPE_A
#define CVQPG_WPE_A 3
#endif
#ifndef CVQPG_PG_RECOMPUTE
#define CVQPG_PG_RECOMPUTE 0
#endif
#ifndef CVQPG_PG_RECOMPUTE_A
#define CVQPG_PG_RECOMPUTE_A 0
#endif
constexpr int ROW_NT = CVQPG_ROW_NT;
constexpr int ROW_LCAP = 8704;   // longest row transform (complex elements)
constexpr int ROW_LCAP_SMALL = 4608;
constexpr int COL_NT = CVQPG_COL_NT;
constexpr int COL_LCAP = 6144;   // column length x columns per workgroup
constexpr int COL_PER = COL_LCAP / COL_NT;

// what a column-pass launch helper does: build the plan's stage tables, run the pass
constexpr int COLS_TABLES = 1, COLS_RUN = 2;

// whether a pass-mask call (re)builds the workspace's transform tables: pass 1 or an
// explicit CVQPG_BLASM_TABLES does, unless CVQPG_BLASM_TABLES_KEPT says they are current
static bool build_tables(int passes) {
  return (passes & (1 | CVQPG_BLASM_TABLES)) && !(passes & CVQPG_BLASM_TABLES_KEPT);
}

struct Optics {
  int C, D, H, W, pH, pW, nu, nv;
  int pad_h, pad_w;      // hologram offset inside the padded field
  int zr0, zc0;          // offset of the padded field on the 2x grid (odak zero_pad)
  int or0, oc0, oh, ow;  // output crop on the 2x grid
  int gr0;               // first 2x-grid row of the hologram (adjoint output rows)
  int cg;                // columns per column workgroup
  int zpair;             // adjoint row spectra Z stored as row pairs (see z_at)
  float norm;            // 1 / (nu nv)
  float fr0, frs;        // centred row-frequency grid  f = fr0 + i frs  (1/m)
  float fc0, fcs;        // centred column-frequency grid
  float lam2[3];         // lambda^2
  float zl[MAXD][3];     // z / lambda (revolutions of the carrier)
  float2 carrier[MAXD][3];  // exp(i 2 pi z / lambda), reduced in FP64 on the host
  float fxmax[MAXD][3];  // band limit on |FX| (column frequency)
  float fymax[MAXD][3];  // band limit on |FY| (row frequency)
  int ap_open;
  int band_active;       // 0 when no band limit and no evanescence can occur on the grid and
                         // every residual phase (z / lambda) g is within 250 revolutions
  float ap_r2, apr0, aprs, apc0, apcs;
  const float* ap_tab;   // custom aperture [nu][nv] on the centred grid, or NULL
};

struct BlasmWS {
  float2* tw_r;   // two-level twiddle table of the column transform's Stockham length
  float2* tw_c;   // two-level twiddle table of the row transform's Stockham length
  float2* ch_r;   // Bluestein chirp and kernel spectrum of the column length (nu, m_r
  float2* bh_r;   //   entries; empty when nu factors into the supported radices)
  float2* ch_c;   // the same for the row length nv
  float2* bh_c;
  float2* tw_cs;  // stage tables of a compile-time column plan (8 nu entries)
  float2* X;      // [C][pH][nv]   (forward pass 1 output; adjoint pass 2 output [C][H][nv])
  float2* Y;      // [D][C][oh][nv] (forward pass 2 output; adjoint pass 3 output)
  size_t bytes;
};

static BlasmWS carve_blasm(void* base, int C, int D, int H, int pH, int pW, int oh) {
  Carver cv(base);
  BlasmWS w;
  const size_t nu = 2 * (size_t)pH, nv = 2 * (size_t)pW;
  const int mr = fft::stockham_len((int)nu), mc = fft::stockham_len((int)nv);
  w.tw_r = cv.take<float2>(fft::twiddle_entries(mr));
  w.tw_c = cv.take<float2>(fft::twiddle_entries(mc));
  w.ch_r = cv.take<float2>(mr != (int)nu ? nu : 0);
  w.bh_r = cv.take<float2>(mr != (int)nu ? mr : 0);
  w.ch_c = cv.take<float2>(mc != (int)nv ? nv : 0);
  w.bh_c = cv.take<float2>(mc != (int)nv ? mc : 0);
  w.tw_cs = cv.take<float2>(8 * nu);
  const size_t xr = (size_t)(pH > H ? pH : H);
  w.X = cv.take<float2>((size_t)C * xr * nv);
  w.Y = cv.take<float2>((size_t)D * C * oh * nv);
  w.bytes = cv.bytes();
  return w;
}

// two-level twiddle table of length n: W_n^{64 h} for h < nh, then W_n^l for l < 64
__device__ __forceinline__ void twiddle_entry(float2* __restrict__ tw, int n, int i) {
  const int nh = (n + fft::TW_LO - 1) / fft::TW_LO;
  if (i >= nh + fft::TW_LO) return;
  const long m = i < nh ? (long)i * fft::TW_LO : (long)(i - nh);
  double s, c;
  sincospi(2.0 * (double)m / (double)n, &s, &c);
  tw[i] = make_float2((float)c, (float)-s);
}

// the tables of both transform lengths (rows nv, columns nu) in one launch; blockIdx.y
// picks the table
__global__ void k_twiddles(float2* __restrict__ tw_r, int nu, float2* __restrict__ tw_c, int nv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.y == 0) twiddle_entry(tw_r, nu, i);
  else twiddle_entry(tw_c, nv, i);
}

// Bluestein tables of length n with Stockham length m (see fft.h), in FP64:
//   chirp[k] = exp(-i pi (k^2 mod 2n) / n),  k < n,
//   bhat[k]  = (1/m) sum_{|j| < n} exp(+i pi j^2 / n) exp(-2 pi i jk / m)
//            = (1/m) (1 + 2 sum_{0 < j < n} exp(+i pi j^2 / n) cos(2 pi jk / m)),  k < m.
// One workgroup per k (the sum over j is split over its threads).  grid m, 256 threads
__global__ __launch_bounds__(256) void k_bluestein_tables(int n, int m, float2* __restrict__ chirp,
                                                          float2* __restrict__ bhat) {
  __shared__ double sr[256], si[256];
  const int k = blockIdx.x, t = threadIdx.x;
  const long two_n = 2L * n;
  double ar = 0.0, ai = 0.0;
  for (int j = 1 + t; j < n; j += 256) {
    double sn, cs;
    sincospi((double)(((long)j * j) % two_n) / (double)n, &sn, &cs);
    const double w = cospi(2.0 * (double)(((long)j * k) % m) / (double)m);
    ar += cs * w;
    ai += sn * w;
  }
  sr[t] = ar;
  si[t] = ai;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off) {
      sr[t] += sr[t + off];
      si[t] += si[t + off];
    }
    __syncthreads();
  }
  if (t == 0) {
    bhat[k] = make_float2((float)((1.0 + 2.0 * sr[0]) / m), (float)(2.0 * si[0] / m));
    if (k < n) {
      double sn, cs;
      sincospi((double)(((long)k * k) % two_n) / (double)n, &sn, &cs);
      chirp[k] = make_float2((float)cs, (float)-sn);
    }
  }
}

// A table of N (compile-time) entries copied from global memory into LDS by NT threads.
// load() issues every load of the thread before store() writes any of them, so the copy
// costs one memory latency instead of one per round of NT entries, and a caller can issue
// its own loads in between; the caller's next barrier publishes the table.
template <int N, int NT>
struct TableCopy {
  static constexpr int R = N > 0 ? (N + NT - 1) / NT : 0;
  float2 v[R > 0 ? R : 1];
  __device__ __forceinline__ void load(const float2* __restrict__ src) {
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[min((int)threadIdx.x + r * NT, N - 1)];
  }
  __device__ __forceinline__ void store(float2* __restrict__ dst) const {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if ((int)threadIdx.x + r * NT < N) dst[threadIdx.x + r * NT] = v[r];
  }
};

// copy the plan's twiddle table into LDS; the caller's next barrier publishes it
__device__ __forceinline__ const float2* stage_twiddles(const fft::Plan& p, float2* stw) {
  const int ne = p.nh + fft::TW_LO;
  for (int i = threadIdx.x; i < ne; i += blockDim.x) stw[i] = p.tw[i];
  return stw;
}

// ---------------------------------------------------------------------------------
// transfer function
// ---------------------------------------------------------------------------------
// Transfer function.  With x = lambda^2 (FX^2 + FY^2),
//   2 pi z sqrt(1/lambda^2 - FX^2 - FY^2) = 2 pi z/lambda - 2 pi (z/lambda) g(x),
//   g(x) = 1 - sqrt(1 - x) = x / (1 + sqrt(1 - x)).
// The first term is a constant carrier phase per (plane, channel); it is reduced modulo
// one revolution in FP64 on the host and applied to the stored plane field.  The second
// term is at most a few revolutions on the grids used here, so FP32 evaluates it to
// ~1e-6 revolution (an FP32 evaluation of the full phase of ~1e4 revolutions would be off
// by ~1e-3 revolution).  Per spectral point the column passes keep only g (negative: no
// transmission, evanescent or outside the aperture).
__device__ __forceinline__ void tf_freqs(const Optics& o, int kr, int kc, int& ir, int& ic,
                                         float& fy, float& fx) {
  // natural FFT index -> centred grid index (the ifftshift of utils/propagator.py:63)
  ir = kr + (o.nu >> 1) >= o.nu ? kr + (o.nu >> 1) - o.nu : kr + (o.nu >> 1);
  ic = kc + (o.nv >> 1) >= o.nv ? kc + (o.nv >> 1) - o.nv : kc + (o.nv >> 1);
  fy = fmaf((float)ir, o.frs, o.fr0);
  fx = fmaf((float)ic, o.fcs, o.fc0);
}

__device__ __forceinline__ float tf_point(const Optics& o, int c, int kr, int kc) {
  int ir, ic;
  float fy, fx;
  tf_freqs(o, kr, kc, ir, ic, fy, fx);
  const float x = o.lam2[c] * fmaf(fx, fx, fy * fy);
  // g is a small number (x << 1 on these grids) whose product with z / lambda is a few
  // revolutions, so the hardware square root and reciprocal (1 ulp) are ample
  float g = x <= 1.0f ? x * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_sqrtf(1.0f - x)) : -1.0f;
  if (!o.ap_open) {
    const float X = fmaf((float)ir, o.aprs, o.apr0), Y = fmaf((float)ic, o.apcs, o.apc0);
    if (!(X * X + Y * Y < o.ap_r2)) g = -1.0f;
  }
  return g;
}

// H_d / carrier at a point: exp(-i 2 pi (z/lambda) g) inside the band, 0 outside.
// BAND = false compiles the point without the band-limit test and hands the residual phase
// to the sine units unreduced (they take revolutions over [-256, 256]); it runs only when no
// band limit and no evanescence can occur on the grid and every residual phase lies within
// 250 revolutions (Optics::band_active == 0).
template <bool BAND = true>
__device__ __forceinline__ float2 tf_value(const Optics& o, float g, int c, int d, int kr, int kc) {
  if (g < 0.0f) return make_float2(0.f, 0.f);
  if (BAND && o.band_active) {
    int ir, ic;
    float fy, fx;
    tf_freqs(o, kr, kc, ir, ic, fy, fx);
    if (!(fabsf(fx) < o.fxmax[d][c]) || !(fabsf(fy) < o.fymax[d][c])) return make_float2(0.f, 0.f);
  }
  const float ph = -o.zl[d][c] * g;
  const float fr = BAND ? __builtin_amdgcn_fractf(ph) : ph;
  return make_float2(__builtin_amdgcn_cosf(fr), __builtin_amdgcn_sinf(fr));
}

// Factor a^2 of a custom aperture a at the spectral point (kr, kc): the reference's
// custom() multiplies the shifted spectrum and the transfer function by the aperture
// (utils/propagator.py:59-61), and a is real, so the forward spectrum and, in the
// adjoint, the plane sum are multiplied by a^2 once for every plane.
__device__ __forceinline__ float ap_square(const Optics& o, int kr, int kc) {
  int ir, ic;
  float fy, fx;
  tf_freqs(o, kr, kc, ir, ic, fy, fx);
  const float a = o.ap_tab[(size_t)ir * o.nv + ic];
  return a * a;
}

// Element (row r, column j) of plane-channel dc of the adjoint row spectra Z
// [D C][oh][nv]: row-major, or with o.zpair as row pairs, rows 2p and 2p + 1 of a column
// side by side, so that the prefetching column kernel (one column per workgroup) loads
// both with one 16-byte access instead of two separate 8-byte pieces 8 nv bytes apart
// (oh is even).
__device__ __forceinline__ size_t z_at(const Optics& o, int dc, int r, int j) {
  const size_t base = (size_t)dc * o.oh * o.nv;
  return o.zpair ? base + ((size_t)(r >> 1) * o.nv + j) * 2 + (r & 1)
                 : base + (size_t)r * o.nv + j;
}

// Row spectra (X, Z: hundreds of MB, read once by the next pass) are written with
// non-temporal stores; whole rows are written, so every line is complete.  The column
// passes' 8- and 16-byte pieces of a row keep ordinary stores (merged in L2).
#ifndef CVQPG_NT_STORES
#define CVQPG_NT_STORES 1
#endif
__device__ __forceinline__ void stream_store(float2* p, float2 v) {
#if CVQPG_NT_STORES
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
#else
  *p = v;
#endif
}

// ---------------------------------------------------------------------------------
// row passes
// ---------------------------------------------------------------------------------
// F1: padded/tapered/glued row -> FFT along the 2x row.  grid (pH, C)
template <class F>
__global__ __launch_bounds__(ROW_NT) void k_f1_rows(const float2* __restrict__ field,
                                                    Optics o, fft::Plan plan,
                                                    float2* __restrict__ X) {
  extern __shared__ float2 buf[];
  const float2* stw = stage_twiddles(plan, buf + fft::lds_elems(plan.len));
  const int r = blockIdx.x, c = blockIdx.y;
  const int sr = min(max(r - o.pad_h, 0), o.H - 1);
  const float wy = taper(r, o.pH, o.pad_h, o.H);
  const float2* src = field + ((size_t)c * o.H + sr) * o.W;
  for (int j = threadIdx.x; j < o.nv; j += ROW_NT) {
    const int q = j - o.zc0;
    float2 u = make_float2(0.f, 0.f);
    if (q >= 0 && q < o.pW) {
      const int sc = min(max(q - o.pad_w, 0), o.W - 1);
      const float w = wy * taper(q, o.pW, o.pad_w, o.W);
      const float2 h = src[sc];
      u = glue(make_float2(h.x * w, h.y * w));
    }
    buf[fft::lidx(j)] = u;
  }
  __syncthreads();
  F::template exec<1, 1, ROW_NT>(buf, plan, stw);
  float2* dst = X + ((size_t)c * o.pH + r) * o.nv;
  for (int j = threadIdx.x; j < o.nv; j += ROW_NT) dst[j] = buf[fft::lidx(j)];
}

// F1 for the propagator interface: u = amplitude e^{i phase} on [C, pH, pW].  grid (pH, C)
template <class F>
__global__ __launch_bounds__(ROW_NT) void k_f1_rows_polar(const float* __restrict__ phase,
                                                          const float* __restrict__ amp,
                                                          Optics o, fft::Plan plan,
                                                          float2* __restrict__ X) {
  extern __shared__ float2 buf[];
  const float2* stw = stage_twiddles(plan, buf + fft::lds_elems(plan.len));
  const int r = blockIdx.x, c = blockIdx.y;
  const size_t base = ((size_t)c * o.pH + r) * o.pW;
  for (int j = threadIdx.x; j < o.nv; j += ROW_NT) {
    const int q = j - o.zc0;
    float2 u = make_float2(0.f, 0.f);
    if (q >= 0 && q < o.pW) {
      const float a = amp[base + q];
      float s, co;
      sincosf(phase[base + q], &s, &co);
      u = make_float2(a * co, a * s);
    }
    buf[fft::lidx(j)] = u;
  }
  __syncthreads();
  F::template exec<1, 1, ROW_NT>(buf, plan, stw);
  float2* dst = X + ((size_t)c * o.pH + r) * o.nv;
  for (int j = threadIdx.x; j < o.nv; j += ROW_NT) dst[j] = buf[fft::lidx(j)];
}

// F3: inverse FFT along the row, crop, field and intensity.  grid (oh, D*C)
template <class F>
__global__ __launch_bounds__(ROW_NT) void k_f3_rows(const float2* __restrict__ Y, Optics o,
                                                    fft::Plan plan, float2* __restrict__ out,
                                                    float* __restrict__ inten) {
  extern __shared__ float2 buf[];
  const float2* stw = stage_twiddles(plan, buf + fft::lds_elems(plan.len));
  const int r = blockIdx.x, dc = blockIdx.y;
  const float2* src = Y + ((size_t)dc * o.oh + r) * o.nv;
  for (int j = threadIdx.x; j < o.nv; j += ROW_NT) buf[fft::lidx(j)] = src[j];
  __syncthreads();
  F::template exec<-1, 1, ROW_NT>(buf, plan, stw);
  const size_t ob = ((size_t)dc * o.oh + r) * o.ow;
  const float2 cr = o.carrier[dc / o.C][dc % o.C];
  for (int x = threadIdx.x; x < o.ow; x += ROW_NT) {
    const float2 v = buf[fft::lidx(o.oc0 + x)];
    const float2 f = make_float2(v.x * o.norm, v.y * o.norm);
    if (out) out[ob + x] = fft::cmul(f, cr);
    inten[ob + x] = f.x * f.x + f.y * f.y;
  }
}

// A3: adjoint of F3.  g_out = 2 dL/dI * out (times the 1/(nu nv) of the forward), placed in
// the cropped columns of a zero 2x row, forward FFT.  grid (oh, D*C)
template <class F>
__global__ __launch_bounds__(ROW_NT) void k_a3_rows(const float* __restrict__ grad_inten,
                                                    const float2* __restrict__ out, Optics o,
                                                    fft::Plan plan, float2* __restrict__ Z) {
  extern __shared__ float2 buf[];
  const float2* stw = stage_twiddles(plan, buf + fft::lds_elems(plan.len));
  const int r = blockIdx.x, dc = blockIdx.y;
  const size_t ob = ((size_t)dc * o.oh + r) * o.ow;
  const float2 cr = o.carrier[dc / o.C][dc % o.C];
  const float2 crc = make_float2(cr.x, -cr.y);
  for (int j = threadIdx.x; j < o.nv; j += ROW_NT) {
    const int x = j - o.oc0;
    float2 g = make_float2(0.f, 0.f);
    if (x >= 0 && x < o.ow) {
      const float s = 2.0f * grad_inten[ob + x] * o.norm;
      const float2 f = fft::cmul(out[ob + x], crc);
      g = make_float2(s * f.x, s * f.y);
    }
    buf[fft::lidx(j)] = g;
  }
  __syncthreads();
  F::template exec<1, 1, ROW_NT>(buf, plan, stw);
  for (int j = threadIdx.x; j < o.nv; j += ROW_NT) Z[z_at(o, dc, r, j)] = buf[fft::lidx(j)];
}

// A1: adjoint of F1 on the hologram rows: unnormalised inverse FFT along the row, the
// hologram columns, adjoint of the glue.  grid (H, C)
template <class F>
__global__ __launch_bounds__(ROW_NT) void k_a1_rows(const float2* __restrict__ G,
                                                    const float2* __restrict__ field, Optics o,
                                                    fft::Plan plan,
                                                    float2* __restrict__ grad_field) {
  extern __shared__ float2 buf[];
  const float2* stw = stage_twiddles(plan, buf + fft::lds_elems(plan.len));
  const int y = blockIdx.x, c = blockIdx.y;
  const float2* src = G + ((size_t)c * o.H + y) * o.nv;
  for (int j = threadIdx.x; j < o.nv; j += ROW_NT) buf[fft::lidx(j)] = src[j];
  __syncthreads();
  F::template exec<-1, 1, ROW_NT>(buf, plan, stw);
  const size_t fb = ((size_t)c * o.H + y) * o.W;
  const int c0 = o.zc0 + o.pad_w;
  for (int x = threadIdx.x; x < o.W; x += ROW_NT)
    grad_field[fb + x] = glue_adjoint(field[fb + x], buf[fft::lidx(c0 + x)]);
}

// A1 for the propagator interface (adjoint of k_f1_rows_polar): unnormalised inverse FFT
// along the row, then, with g = dL/dRe u + i dL/dIm u at u = a e^{i phi},
//   dL/dphi = Re(conj(g) i u) = g.y u.x - g.x u.y,   dL/da = g.x cos phi + g.y sin phi.
// grid (pH, C)
template <class F>
__global__ __launch_bounds__(ROW_NT) void k_a1_rows_polar(const float2* __restrict__ G,
                                                          const float* __restrict__ phase,
                                                          const float* __restrict__ amp,
                                                          Optics o, fft::Plan plan,
                                                          float* __restrict__ grad_phase,
                                                          float* __restrict__ grad_amp) {
  extern __shared__ float2 buf[];
  const float2* stw = stage_twiddles(plan, buf + fft::lds_elems(plan.len));
  const int y = blockIdx.x, c = blockIdx.y;
  const float2* src = G + ((size_t)c * o.H + y) * o.nv;
  for (int j = threadIdx.x; j < o.nv; j += ROW_NT) buf[fft::lidx(j)] = src[j];
  __syncthreads();
  F::template exec<-1, 1, ROW_NT>(buf, plan, stw);
  const size_t base = ((size_t)c * o.pH + y) * o.pW;
  for (int x = threadIdx.x; x < o.pW; x += ROW_NT) {
    const float2 g = buf[fft::lidx(o.zc0 + x)];
    float s, co;
    sincosf(phase[base + x], &s, &co);
    const float a = amp[base + x];
    grad_phase[base + x] = a * (g.y * co - g.x * s);
    grad_amp[base + x] = g.x * co + g.y * s;
  }
}

// ---------------------------------------------------------------------------------
// column passes (CG adjacent columns per workgroup, XCD-aware block order)
// ---------------------------------------------------------------------------------
__device__ __forceinline__ int xcd_block(int hb, int per) { return (hb & 7) * per + (hb >> 3); }

// F2: column FFT, per plane transfer function, inverse column FFT, crop rows.
// grid (8 * per, C)
template <int CG, class F>
__global__ __launch_bounds__(COL_NT) void k_f2_cols(const float2* __restrict__ X, Optics o,
                                                    fft::Plan plan_fwd, int per,
                                                    float2* __restrict__ Y) {
  extern __shared__ float2 buf[];
  const int lb = xcd_block(blockIdx.x, per);
  const int col0 = lb * CG;
  if (col0 >= o.nv) return;
  const int c = blockIdx.y;
  const int n = o.nu;
  const int total = n * CG;
  const float2* stw = stage_twiddles(plan_fwd, buf + fft::lds_elems(plan_fwd.len * CG));
  for (int e = threadIdx.x; e < total; e += COL_NT) {
    const int i = e / CG, g = e % CG;
    const int r = i - o.zr0;
    buf[fft::lidx(e)] = (r >= 0 && r < o.pH) ? X[((size_t)c * o.pH + r) * o.nv + col0 + g]
                                  : make_float2(0.f, 0.f);
  }
  __syncthreads();
  F::template exec<1, CG, COL_NT>(buf, plan_fwd, stw);
  float2 S[COL_PER];
  float P[COL_PER];
#pragma unroll
  for (int b = 0; b < COL_PER; ++b) {
    const int e = threadIdx.x + b * COL_NT;
    if (e < total) {
      S[b] = buf[fft::lidx(e)];
      P[b] = tf_point(o, c, e / CG, col0 + e % CG);
      if (o.ap_tab) S[b] = fft::cscale(S[b], ap_square(o, e / CG, col0 + e % CG));
    }
  }
  for (int d = 0; d < o.D; ++d) {
    __syncthreads();
#pragma unroll
    for (int b = 0; b < COL_PER; ++b) {
      const int e = threadIdx.x + b * COL_NT;
      if (e < total) buf[fft::lidx(e)] = fft::cmul(S[b], tf_value(o, P[b], c, d, e / CG, col0 + e % CG));
    }
    __syncthreads();
    F::template exec<-1, CG, COL_NT>(buf, plan_fwd, stw);
    float2* dst = Y + ((size_t)d * o.C + c) * o.oh * o.nv;
    for (int e = threadIdx.x; e < o.oh * CG; e += COL_NT) {
      const int i = e / CG, g = e % CG;
      dst[(size_t)i * o.nv + col0 + g] = buf[fft::lidx((o.or0 + i) * CG + g)];
    }
  }
}

// A2: for every plane, column FFT of the adjoint rows, multiply by conj(H_d), sum; then
// unnormalised inverse column FFT and keep the hologram rows.  grid (8 * per, C)
template <int CG, class F>
__global__ __launch_bounds__(COL_NT) void k_a2_cols(const float2* __restrict__ Z, Optics o,
                                                    fft::Plan plan, int per,
                                                    float2* __restrict__ G) {
  extern __shared__ float2 buf[];
  const int lb = xcd_block(blockIdx.x, per);
  const int col0 = lb * CG;
  if (col0 >= o.nv) return;
  const int c = blockIdx.y;
  const int total = o.nu * CG;
  const float2* stw = stage_twiddles(plan, buf + fft::lds_elems(plan.len * CG));
  float2 A[COL_PER];
  float P[COL_PER];
#pragma unroll
  for (int b = 0; b < COL_PER; ++b) {
    const int e = threadIdx.x + b * COL_NT;
    A[b] = make_float2(0.f, 0.f);
    if (e < total) P[b] = tf_point(o, c, e / CG, col0 + e % CG);
  }
  for (int d = 0; d < o.D; ++d) {
    const float2* src = Z + ((size_t)d * o.C + c) * o.oh * o.nv;
    for (int e = threadIdx.x; e < total; e += COL_NT) {
      const int i = e / CG, g = e % CG;
      const int r = i - o.or0;
      buf[fft::lidx(e)] = (r >= 0 && r < o.oh) ? src[(size_t)r * o.nv + col0 + g] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    F::template exec<1, CG, COL_NT>(buf, plan, stw);
#pragma unroll
    for (int b = 0; b < COL_PER; ++b) {
      const int e = threadIdx.x + b * COL_NT;
      if (e < total) {
        const float2 h = tf_value(o, P[b], c, d, e / CG, col0 + e % CG);
        const float2 v = buf[fft::lidx(e)];
        A[b].x += v.x * h.x + v.y * h.y;  // v * conj(h)
        A[b].y += v.y * h.x - v.x * h.y;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int b = 0; b < COL_PER; ++b) {
    const int e = threadIdx.x + b * COL_NT;
    if (e < total)
      buf[fft::lidx(e)] = o.ap_tab ? fft::cscale(A[b], ap_square(o, e / CG, col0 + e % CG)) : A[b];
  }
  __syncthreads();
  F::template exec<-1, CG, COL_NT>(buf, plan, stw);
  float2* dst = G + (size_t)c * o.H * o.nv;
  for (int e = threadIdx.x; e < o.H * CG; e += COL_NT) {
    const int i = e / CG, g = e % CG;
    dst[(size_t)i * o.nv + col0 + g] = buf[fft::lidx((o.gr0 + i) * CG + g)];
  }
}

// ---------------------------------------------------------------------------------
// compile-time pipelines for the transform lengths of the standard workloads
// ---------------------------------------------------------------------------------
// The first butterfly stage reads its inputs straight from HBM (or builds them from the
// raster output) into registers and the last stage writes its outputs straight to HBM,
// so a transform of L stages makes L - 1 LDS round trips instead of L + 1.  On the 2x
// grid only positions [N/4, 3N/4) of a padded row or column are non-zero; with a first
// radix divisible by 4 that is exactly r in [R/4, 3R/4) of every first-stage butterfly,
// so the zero half is neither loaded nor stored.  The column kernels keep the forward
// spectrum in the last-stage registers, which are the first-stage inputs of the inverse
// transform (the inverse runs the radix list reversed).

// RowPlan<NT, Fwd, Inv>: row transform of length Fwd::N on NT threads; ColPlan<CG, NT,
// WPE_F, WPE_A, Fwd, Inv>: column transform of CG columns on NT threads, the forward and
// adjoint kernels compiled for at least WPE_F / WPE_A waves per SIMD.  Inv runs Fwd's
// radices reversed (its first stage consumes the forward's last-stage registers); the
// buffer layouts are the ones a 64-bank conflict model picks for each stage boundary.
template <class F, class I>
struct PlanPair {
  using Fwd = F;
  using Inv = I;
  static_assert(F::FIRST == I::LAST && F::LAST == I::FIRST, "inverse must reverse the forward");
  static constexpr int n = F::n_;
};
template <int NT, class F, class I>
struct RowPlan : PlanPair<F, I> {
  static constexpr int nt = NT;
  static constexpr int lds = (F::lds_elems(1) > I::lds_elems(1) ? F::lds_elems(1) : I::lds_elems(1));
};
template <int CG, int NT, int WPE_F, int WPE_A, class F, class I>
struct ColPlan : PlanPair<F, I> {
  static constexpr int cg = CG, nt = NT, wpe_f = WPE_F, wpe_a = WPE_A;
  static constexpr int lds = (F::lds_elems(CG) > I::lds_elems(CG) ? F::lds_elems(CG) : I::lds_elems(CG));
};
using fft::Pipe;
using fft::Rad;
using fft::Lays;
template <int P, int Q>
using Ly = fft::Lay<P, Q>;
using Ly0 = fft::Lay<0, 0>;

template <class Pp>
__device__ __forceinline__ bool in_mid_half(int r) {
  return r >= Pp::FIRST / 4 && r < 3 * Pp::FIRST / 4;
}

// F1 (see k_f1_rows) with a compile-time plan.  grid (pH, C)
template <class P>
__global__ __launch_bounds__(P::nt) void k_f1_rows_fx(const float2* __restrict__ field,
                                                      Optics o, fft::Plan plan,
                                                      float2* __restrict__ X) {
  constexpr int N = P::n, NT = P::nt;
  using Fw = typename P::Fwd;
  using G1 = typename Fw::template G1<1, NT>;
  using GL = typename Fw::template GL<1, NT>;
  static_assert(Fw::FIRST % 4 == 0, "first radix must be divisible by 4");
  extern __shared__ float2 buf[];
  const float2* stw = stage_twiddles(plan, buf + P::lds);
  const int r = blockIdx.x, c = blockIdx.y;
  const int sr = min(max(r - o.pad_h, 0), o.H - 1);
  const float wy = taper(r, o.pH, o.pad_h, o.H);
  const float2* src = field + ((size_t)c * o.H + sr) * o.W;
  {
    float2 v[G1::MAXB][Fw::FIRST];
#pragma unroll
    for (int b = 0; b < G1::MAXB; ++b) {
      if (!G1::live(b)) continue;
      int j, col;
      G1::item(b, j, col);
      // all loads of the row first (a load guarded together with its use would wait for
      // each in turn); columns outside the padded field read the clamped edge
      float2 hv[Fw::FIRST];
#pragma unroll
      for (int k = 0; k < Fw::FIRST; ++k) {
        if (!in_mid_half<Fw>(k)) continue;
        const int q = j + k * G1::CNT - o.zc0;
        hv[k] = src[min(max(q - o.pad_w, 0), o.W - 1)];
      }
#pragma unroll
      for (int k = 0; k < Fw::FIRST; ++k) {
        float2 u = make_float2(0.f, 0.f);
        if (in_mid_half<Fw>(k)) {
          const int q = j + k * G1::CNT - o.zc0;
          if ((unsigned)q < (unsigned)o.pW) {
            const float w = wy * taper(q, o.pW, o.pad_w, o.W);
            u = glue(make_float2(hv[k].x * w, hv[k].y * w));
          }
        }
        v[b][k] = u;
      }
    }
    Fw::template first<1, 1, NT>(buf, v);
  }
  __syncthreads();
  Fw::template middle<1, 1, NT>(buf, stw, plan.nh);
  float2 v[GL::MAXB][Fw::LAST];
  Fw::template last<1, 1, NT>(buf, stw, plan.nh, v);
  float2* dst = X + ((size_t)c * o.pH + r) * o.nv;
#pragma unroll
  for (int b = 0; b < GL::MAXB; ++b) {
    if (!GL::live(b)) continue;
    int j, col;
    GL::item(b, j, col);
#pragma unroll
    for (int k = 0; k < Fw::LAST; ++k) stream_store(dst + j + k * GL::CNT, v[b][k]);
  }
}

// F3 (see k_f3_rows) with a compile-time plan.  grid (oh, D*C)
template <class P>
__global__ __launch_bounds__(P::nt) void k_f3_rows_fx(const float2* __restrict__ Y, Optics o,
                                                      fft::Plan plan, float2* __restrict__ out,
                                                      float* __restrict__ inten) {
  constexpr int N = P::n, NT = P::nt;
  using Iv = typename P::Inv;
  using G1 = typename Iv::template G1<1, NT>;
  using GL = typename Iv::template GL<1, NT>;
  extern __shared__ float2 buf[];
  const float2* stw = stage_twiddles(plan, buf + P::lds);
  const int r = blockIdx.x, dc = blockIdx.y;
  const float2* src = Y + ((size_t)dc * o.oh + r) * o.nv;
  {
    float2 v[G1::MAXB][Iv::FIRST];
#pragma unroll
    for (int b = 0; b < G1::MAXB; ++b) {
      if (!G1::live(b)) continue;
      int j, col;
      G1::item(b, j, col);
#pragma unroll
      for (int k = 0; k < Iv::FIRST; ++k) v[b][k] = src[j + k * G1::CNT];
    }
    Iv::template first<-1, 1, NT>(buf, v);
  }
  __syncthreads();
  Iv::template middle<-1, 1, NT>(buf, stw, plan.nh);
  float2 v[GL::MAXB][Iv::LAST];
  Iv::template last<-1, 1, NT>(buf, stw, plan.nh, v);
  const size_t ob = ((size_t)dc * o.oh + r) * o.ow;
  const float2 cr = o.carrier[dc / o.C][dc % o.C];
#pragma unroll
  for (int b = 0; b < GL::MAXB; ++b) {
    if (!GL::live(b)) continue;
    int j, col;
    GL::item(b, j, col);
#pragma unroll
    for (int k = Iv::LAST / 4; k < 3 * Iv::LAST / 4; ++k) {   // the crop lies in [N/4, 3N/4)
      const int x = j + k * GL::CNT - o.oc0;
      if ((unsigned)x < (unsigned)o.ow) {
        const float2 f = make_float2(v[b][k].x * o.norm, v[b][k].y * o.norm);
        if (out) out[ob + x] = fft::cmul(f, cr);
        inten[ob + x] = f.x * f.x + f.y * f.y;
      }
    }
  }
}

// A3 (see k_a3_rows) with a compile-time plan: the row spectra of rows r0 + g, g < CG, of
// plane-channel dc (CG interleaved sequences in buf) into the last-stage registers v
// (synchronised on entry; ends with buf still being read)
template <class P, int CG, int NT>
__device__ __forceinline__ void a3_rows_fx(
    const float* __restrict__ grad_inten, const float2* __restrict__ out, const Optics& o,
    const fft::Plan& plan, const float2* stw, float2* buf, int r0, int dc,
    float2 (&v)[P::Fwd::template GL<CG, NT>::MAXB][P::Fwd::LAST]) {
  using Fw = typename P::Fwd;
  using G1 = typename Fw::template G1<CG, NT>;
  static_assert(Fw::FIRST % 4 == 0, "first radix must be divisible by 4");
  const float2 cr = o.carrier[dc / o.C][dc % o.C];
  const float2 crc = make_float2(cr.x, -cr.y);
  {
    float2 u[G1::MAXB][Fw::FIRST];
#pragma unroll
    for (int b = 0; b < G1::MAXB; ++b) {
      if (!G1::live(b)) continue;
      int j, g;
      G1::item(b, j, g);
      const size_t ob = ((size_t)dc * o.oh + r0 + g) * o.ow;
      // every load of the row is issued before any is used (a load guarded together with
      // its use would wait for each in turn); out-of-crop elements read element 0
      float gi[Fw::FIRST];
      float2 fo[Fw::FIRST];
#pragma unroll
      for (int k = 0; k < Fw::FIRST; ++k) {
        if (!in_mid_half<Fw>(k)) continue;
        const int x = j + k * G1::CNT - o.oc0;
        const size_t e = ob + ((unsigned)x < (unsigned)o.ow ? x : 0);
        gi[k] = grad_inten[e];
        fo[k] = out[e];
      }
#pragma unroll
      for (int k = 0; k < Fw::FIRST; ++k) {
        float2 q = make_float2(0.f, 0.f);
        if (in_mid_half<Fw>(k)) {
          const int x = j + k * G1::CNT - o.oc0;
          if ((unsigned)x < (unsigned)o.ow) {
            const float sc = 2.0f * gi[k] * o.norm;
            const float2 f = fft::cmul(fo[k], crc);
            q = make_float2(sc * f.x, sc * f.y);
          }
        }
        u[b][k] = q;
      }
    }
    Fw::template first<1, CG, NT>(buf, u);
  }
  __syncthreads();
  Fw::template middle<1, CG, NT>(buf, stw, plan.nh);
  Fw::template last<1, CG, NT>(buf, stw, plan.nh, v);
}

// A3 with a compile-time plan, row-major Z: one row per workgroup.  grid (oh, D*C)
template <class P>
__global__ __launch_bounds__(P::nt) void k_a3_rows_fx(const float* __restrict__ grad_inten,
                                                      const float2* __restrict__ out, Optics o,
                                                      fft::Plan plan, float2* __restrict__ Z) {
  constexpr int NT = P::nt;
  using Fw = typename P::Fwd;
  using GL = typename Fw::template GL<1, NT>;
  extern __shared__ float2 buf[];
  const float2* stw = stage_twiddles(plan, buf + P::lds);
  const int r = blockIdx.x, dc = blockIdx.y;
  float2 v[GL::MAXB][Fw::LAST];
  a3_rows_fx<P, 1, NT>(grad_inten, out, o, plan, stw, buf, r, dc, v);
  float2* dst = Z + z_at(o, dc, r, 0);
#pragma unroll
  for (int b = 0; b < GL::MAXB; ++b) {
    if (!GL::live(b)) continue;
    int j, g;
    GL::item(b, j, g);
#pragma unroll
    for (int k = 0; k < Fw::LAST; ++k) stream_store(dst + j + k * GL::CNT, v[b][k]);
  }
}

// A3 with a compile-time plan and Z as row pairs (o.zpair): rows 2p and 2p + 1 per
// workgroup, so every line of Z is written whole.  When twice the plan's threads fit a
// workgroup, the two rows are transformed together as two interleaved sequences
// (2 nt threads; lanes 2m and 2m + 1 hold the two rows of one column, which are adjacent
// in Z, so a wave's stores are contiguous); otherwise one after the other on nt threads,
// the first row's spectrum kept in registers and each column pair stored as one 16-byte
// piece.  grid (oh / 2, D*C)
#ifndef CVQPG_A3_PAIR_THREADS
#define CVQPG_A3_PAIR_THREADS 2   // threads of a two-row workgroup, in units of the plan's nt
#endif
template <class P>
constexpr bool a3_pair_together() {
  return CVQPG_A3_PAIR_THREADS * P::nt <= 1024 &&
         (P::Fwd::lds_elems(2) + fft::twiddle_entries(P::n)) * sizeof(float2) <= 80 * 1024;
}
template <class P>
constexpr int a3_pair_threads() {
  return a3_pair_together<P>() ? CVQPG_A3_PAIR_THREADS * P::nt : P::nt;
}
template <class P>
constexpr int a3_pair_lds() { return a3_pair_together<P>() ? P::Fwd::lds_elems(2) : P::lds; }

template <class P>
__global__ __launch_bounds__(a3_pair_threads<P>()) void k_a3_rows_pair_fx(
    const float* __restrict__ grad_inten, const float2* __restrict__ out, Optics o,
    fft::Plan plan, float2* __restrict__ Z) {
  using Fw = typename P::Fwd;
  extern __shared__ float2 buf[];
  const float2* stw = stage_twiddles(plan, buf + a3_pair_lds<P>());
  const int r = 2 * blockIdx.x, dc = blockIdx.y;
  if constexpr (a3_pair_together<P>()) {
    constexpr int NT = a3_pair_threads<P>();
    using GL = typename Fw::template GL<2, NT>;
    float2 v[GL::MAXB][Fw::LAST];
    a3_rows_fx<P, 2, NT>(grad_inten, out, o, plan, stw, buf, r, dc, v);
#pragma unroll
    for (int b = 0; b < GL::MAXB; ++b) {
      if (!GL::live(b)) continue;
      int j, g;
      GL::item(b, j, g);
      float2* dst = Z + z_at(o, dc, r, j) + g;   // row r + g, column j
#pragma unroll
      for (int k = 0; k < Fw::LAST; ++k) stream_store(dst + 2 * k * GL::CNT, v[b][k]);
    }
  } else {
    constexpr int NT = P::nt;
    using GL = typename Fw::template GL<1, NT>;
    float2 v[GL::MAXB][Fw::LAST], w[GL::MAXB][Fw::LAST];
    a3_rows_fx<P, 1, NT>(grad_inten, out, o, plan, stw, buf, r, dc, v);
    __syncthreads();   // every read of buf by the first row's last stage is complete
    a3_rows_fx<P, 1, NT>(grad_inten, out, o, plan, stw, buf, r + 1, dc, w);
    typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int b = 0; b < GL::MAXB; ++b) {
      if (!GL::live(b)) continue;
      int j, g;
      GL::item(b, j, g);
      f4* dst = reinterpret_cast<f4*>(Z + z_at(o, dc, r, j));
#pragma unroll
      for (int k = 0; k < Fw::LAST; ++k) {
        const f4 q = {v[b][k].x, v[b][k].y, w[b][k].x, w[b][k].y};
        __builtin_nontemporal_store(q, dst + k * GL::CNT);
      }
    }
  }
}

// A1 (see k_a1_rows) with a compile-time plan.  grid (H, C)
template <class P>
__global__ __launch_bounds__(P::nt) void k_a1_rows_fx(const float2* __restrict__ G,
                                                      const float2* __restrict__ field, Optics o,
                                                      fft::Plan plan,
                                                      float2* __restrict__ grad_field) {
  constexpr int N = P::n, NT = P::nt;
  using Iv = typename P::Inv;
  using G1 = typename Iv::template G1<1, NT>;
  using GL = typename Iv::template GL<1, NT>;
  extern __shared__ float2 buf[];
  const float2* stw = stage_twiddles(plan, buf + P::lds);
  const int y = blockIdx.x, c = blockIdx.y;
  const float2* src = G + ((size_t)c * o.H + y) * o.nv;
  {
    float2 v[G1::MAXB][Iv::FIRST];
#pragma unroll
    for (int b = 0; b < G1::MAXB; ++b) {
      if (!G1::live(b)) continue;
      int j, col;
      G1::item(b, j, col);
#pragma unroll
      for (int k = 0; k < Iv::FIRST; ++k) v[b][k] = src[j + k * G1::CNT];
    }
    Iv::template first<-1, 1, NT>(buf, v);
  }
  __syncthreads();
  Iv::template middle<-1, 1, NT>(buf, stw, plan.nh);
  float2 v[GL::MAXB][Iv::LAST];
  Iv::template last<-1, 1, NT>(buf, stw, plan.nh, v);
  const size_t fb = ((size_t)c * o.H + y) * o.W;
  const int c0 = o.zc0 + o.pad_w;
#pragma unroll
  for (int b = 0; b < GL::MAXB; ++b) {
    if (!GL::live(b)) continue;
    int j, col;
    GL::item(b, j, col);
    // the field values of every output first (a load guarded together with its use would
    // wait for each in turn); outputs outside the hologram read element 0
    float2 fv[Iv::LAST];
#pragma unroll
    for (int k = Iv::LAST / 4; k < 3 * Iv::LAST / 4; ++k) {   // the hologram lies in [N/4, 3N/4)
      const int x = j + k * GL::CNT - c0;
      fv[k] = field[fb + ((unsigned)x < (unsigned)o.W ? x : 0)];
    }
#pragma unroll
    for (int k = Iv::LAST / 4; k < 3 * Iv::LAST / 4; ++k) {
      const int x = j + k * GL::CNT - c0;
      if ((unsigned)x < (unsigned)o.W) grad_field[fb + x] = glue_adjoint(fv[k], v[b][k]);
    }
  }
}

// LDS of a column kernel: data, two-level table, then (16-byte aligned) stage tables
template <class P>
__host__ __device__ constexpr int col_lds_head() {
  return (P::lds + fft::twiddle_entries(P::n) + 1) & ~1;
}

// global stage tables of a column plan: the forward pipe's, then the inverse pipe's
template <class P>
__global__ void k_pipe_tables(float2* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P::Fwd::TABLE) P::Fwd::table_entry(out, i);
  else if (i < P::Fwd::TABLE + P::Inv::TABLE) P::Inv::table_entry(out + P::Fwd::TABLE, i - P::Fwd::TABLE);
}

// F2 (see k_f2_cols) with a compile-time plan.  The per-plane inverse transforms read
// their twiddles from the inverse pipe's stage tables, staged in LDS once per
// workgroup; only the rows of the inverse that can fall in the crop ([N/4, 3N/4)) are
// formed.  grid (8 * per, C)
template <class P, bool BAND>
__global__ __launch_bounds__(P::nt, P::wpe_f) void k_f2_cols_fx(const float2* __restrict__ X, Optics o,
                                                              fft::Plan plan,
                                                              const float2* __restrict__ tabs,
                                                              int per, float2* __restrict__ Y) {
  constexpr int N = P::n, NT = P::nt, CG = P::cg;
  using Fw = typename P::Fwd;
  using Iv = typename P::Inv;
  using F1 = typename Fw::template G1<CG, NT>;
  using FL = typename Fw::template GL<CG, NT>;   // == first stage of the inverse
  using IL = typename Iv::template GL<CG, NT>;
  static_assert(Fw::FIRST % 4 == 0 && Iv::LAST % 4 == 0, "first radix must be divisible by 4");
  static_assert(Fw::LAST == Iv::FIRST, "inverse must start with the forward's last radix");
  extern __shared__ float2 buf[];
  const int lb = xcd_block(blockIdx.x, per);
  const int col0 = lb * CG;
  if (col0 >= o.nv) return;
  const int c = blockIdx.y;
  const float2* stw = stage_twiddles(plan, buf + P::lds);
  float2* tt = buf + col_lds_head<P>();
  TableCopy<Iv::TABLE, NT> tcopy;
  tcopy.load(tabs + Fw::TABLE);   // stored after the column's loads are issued
  // plane bases are uniform; offsets inside a plane (< 2^32 elements) are 32-bit
  const float2* Xc = X + (size_t)c * o.pH * o.nv + col0;
  {
    float2 v[F1::MAXB][Fw::FIRST];
#pragma unroll
    for (int b = 0; b < F1::MAXB; ++b) {
      if (!F1::live(b)) continue;
      int j, col;
      F1::item(b, j, col);
#pragma unroll
      for (int k = 0; k < Fw::FIRST; ++k) {
        float2 u = make_float2(0.f, 0.f);
        if (in_mid_half<Fw>(k)) {
          const int rr = j + k * F1::CNT - o.zr0;
          if ((unsigned)rr < (unsigned)o.pH) u = Xc[(unsigned)(rr * o.nv + col)];
        }
        v[b][k] = u;
      }
    }
    tcopy.store(tt);
    Fw::template first<1, CG, NT>(buf, v);
  }
  __syncthreads();
  Fw::template middle<1, CG, NT>(buf, stw, plan.nh);
  float2 S[FL::MAXB][Fw::LAST];
  Fw::template last<1, CG, NT>(buf, stw, plan.nh, S);
  if (o.ap_tab) {
#pragma unroll
    for (int b = 0; b < FL::MAXB; ++b) {
      if (!FL::live(b)) continue;
      int j, col;
      FL::item(b, j, col);
#pragma unroll
      for (int k = 0; k < Fw::LAST; ++k)
        S[b][k] = fft::cscale(S[b][k], ap_square(o, j + k * FL::CNT, col0 + col));
    }
  }
#if !CVQPG_PG_RECOMPUTE
  float Pg[FL::MAXB][Fw::LAST];
#pragma unroll
  for (int b = 0; b < FL::MAXB; ++b) {
    if (!FL::live(b)) continue;
    int j, col;
    FL::item(b, j, col);
#pragma unroll
    for (int k = 0; k < Fw::LAST; ++k) Pg[b][k] = tf_point(o, c, j + k * FL::CNT, col0 + col);
  }
#endif
  for (int d = 0; d < o.D; ++d) {
    __syncthreads();   // the previous reads of buf are complete
    {
      float2 v[FL::MAXB][Fw::LAST];
#pragma unroll
      for (int b = 0; b < FL::MAXB; ++b) {
        if (!FL::live(b)) continue;
        int j, col;
        FL::item(b, j, col);
#pragma unroll
        for (int k = 0; k < Fw::LAST; ++k) {
#if CVQPG_PG_RECOMPUTE
          const float pg = tf_point(o, c, j + k * FL::CNT, col0 + col);
#else
          const float pg = Pg[b][k];
#endif
          v[b][k] = fft::cmul(S[b][k], tf_value<BAND>(o, pg, c, d, j + k * FL::CNT, col0 + col));
        }
      }
      Iv::template first<-1, CG, NT>(buf, v);
    }
    __syncthreads();
    Iv::template middle<-1, CG, NT, true>(buf, stw, plan.nh, tt);
    float2 u[IL::MAXB][Iv::LAST];
    Iv::template last<-1, CG, NT, true>(buf, stw, plan.nh, u, tt);
    // 32-bit element offsets from Y + col0 that include the plane: the eight row offsets of
    // a thread are then not loop-invariant 64-bit addresses the compiler keeps across the
    // plane loop (they spilled, and reloading a spilled address before a store makes the
    // wave wait for every memory operation in flight, the previous plane's stores
    // included).  Y holds D C oh nv < 2^32 elements (planes <= 16, oh <= 3072, nv <= 8704).
    const unsigned pe = (unsigned)((d * o.C + c) * o.oh) * (unsigned)o.nv;
    float2* Yc = Y + col0;
#pragma unroll
    for (int b = 0; b < IL::MAXB; ++b) {
      if (!IL::live(b)) continue;
      int j, col;
      IL::item(b, j, col);
#pragma unroll
      for (int k = Iv::LAST / 4; k < 3 * Iv::LAST / 4; ++k) {
        const int rr = j + k * IL::CNT - o.or0;
        if ((unsigned)rr < (unsigned)o.oh) Yc[pe + (unsigned)(rr * o.nv + col)] = u[b][k];
      }
    }
  }
}

// A2 (see k_a2_cols) with a compile-time plan.  The per-plane forward transforms read
// their twiddles from the forward pipe's stage tables staged in LDS.  grid (8 * per, C)
template <class P, bool BAND>
__global__ __launch_bounds__(P::nt, P::wpe_a) void k_a2_cols_fx(const float2* __restrict__ Z, Optics o,
                                                              fft::Plan plan,
                                                              const float2* __restrict__ tabs,
                                                              int per, float2* __restrict__ G) {
  constexpr int N = P::n, NT = P::nt, CG = P::cg;
  using Fw = typename P::Fwd;
  using Iv = typename P::Inv;
  using F1 = typename Fw::template G1<CG, NT>;
  using FL = typename Fw::template GL<CG, NT>;
  using IL = typename Iv::template GL<CG, NT>;
  static_assert(Fw::FIRST % 4 == 0 && Iv::LAST % 4 == 0, "first radix must be divisible by 4");
  static_assert(Fw::LAST == Iv::FIRST, "inverse must start with the forward's last radix");
  extern __shared__ float2 buf[];
  const int lb = xcd_block(blockIdx.x, per);
  const int col0 = lb * CG;
  if (col0 >= o.nv) return;
  const int c = blockIdx.y;
  const float2* stw = stage_twiddles(plan, buf + P::lds);
  float2* tt = buf + col_lds_head<P>();
  {
    TableCopy<Fw::TABLE, NT> tcopy;
    tcopy.load(tabs);
    tcopy.store(tt);
  }
  float2 A[FL::MAXB][Fw::LAST];
#if !CVQPG_PG_RECOMPUTE_A
  float Pg[FL::MAXB][Fw::LAST];
#endif
#pragma unroll
  for (int b = 0; b < FL::MAXB; ++b) {
    int j, col;
    FL::item(b, j, col);
#pragma unroll
    for (int k = 0; k < Fw::LAST; ++k) {
      A[b][k] = make_float2(0.f, 0.f);
#if !CVQPG_PG_RECOMPUTE_A
      Pg[b][k] = FL::live(b) ? tf_point(o, c, j + k * FL::CNT, col0 + col) : -1.0f;
#endif
    }
  }
  for (int d = 0; d < o.D; ++d) {
    if (d > 0) __syncthreads();   // the previous plane's reads of buf are complete
    const float2* src = Z + ((size_t)d * o.C + c) * o.oh * o.nv + col0;
    {
      float2 v[F1::MAXB][Fw::FIRST];
#pragma unroll
      for (int b = 0; b < F1::MAXB; ++b) {
        if (!F1::live(b)) continue;
        int j, col;
        F1::item(b, j, col);
#pragma unroll
        for (int k = 0; k < Fw::FIRST; ++k) {
          float2 u = make_float2(0.f, 0.f);
          if (in_mid_half<Fw>(k)) {
            const int rr = j + k * F1::CNT - o.or0;
            if ((unsigned)rr < (unsigned)o.oh) u = src[(unsigned)(rr * o.nv + col)];
          }
          v[b][k] = u;
        }
      }
      Fw::template first<1, CG, NT>(buf, v);
    }
    __syncthreads();
    Fw::template middle<1, CG, NT, true>(buf, stw, plan.nh, tt);
    float2 v[FL::MAXB][Fw::LAST];
    Fw::template last<1, CG, NT, true>(buf, stw, plan.nh, v, tt);
#pragma unroll
    for (int b = 0; b < FL::MAXB; ++b) {
      if (!FL::live(b)) continue;
      int j, col;
      FL::item(b, j, col);
#pragma unroll
      for (int k = 0; k < Fw::LAST; ++k) {
#if CVQPG_PG_RECOMPUTE_A
        const float pg = tf_point(o, c, j + k * FL::CNT, col0 + col);
#else
        const float pg = Pg[b][k];
#endif
        const float2 h = tf_value<BAND>(o, pg, c, d, j + k * FL::CNT, col0 + col);
        A[b][k].x += v[b][k].x * h.x + v[b][k].y * h.y;   // v * conj(h)
        A[b][k].y += v[b][k].y * h.x - v[b][k].x * h.y;
      }
    }
  }
  __syncthreads();
  if (o.ap_tab) {
#pragma unroll
    for (int b = 0; b < FL::MAXB; ++b) {
      if (!FL::live(b)) continue;
      int j, col;
      FL::item(b, j, col);
#pragma unroll
      for (int k = 0; k < Fw::LAST; ++k)
        A[b][k] = fft::cscale(A[b][k], ap_square(o, j + k * FL::CNT, col0 + col));
    }
  }
  Iv::template first<-1, CG, NT>(buf, A);
  __syncthreads();
  Iv::template middle<-1, CG, NT>(buf, stw, plan.nh);
  float2 u[IL::MAXB][Iv::LAST];
  Iv::template last<-1, CG, NT>(buf, stw, plan.nh, u);
  float2* dst = G + (size_t)c * o.H * o.nv + col0;
#pragma unroll
  for (int b = 0; b < IL::MAXB; ++b) {
    if (!IL::live(b)) continue;
    int j, col;
    IL::item(b, j, col);
#pragma unroll
    for (int k = Iv::LAST / 4; k < 3 * Iv::LAST / 4; ++k) {
      const int rr = j + k * IL::CNT - o.gr0;
      if ((unsigned)rr < (unsigned)o.H) dst[(unsigned)(rr * o.nv + col)] = u[b][k];
    }
  }
}

// A2 for the compile-time plans of one column per workgroup (CG = 1), with the next
// plane's column prefetched into LDS while the current plane is transformed.  The plane's
// oh rows of one column are 8-byte pieces 8 nv bytes apart; loaded into registers they
// would have to be waited for before each plane's transform.  Here they are copied with
// direct-to-LDS loads (global_load_lds_dword: lane i of a wave writes word i after the
// wave's LDS base; real and imaginary parts go to two arrays), issued right after the
// plane's first stage has consumed the previous copy, and waited for only at the next
// plane's first barrier; the stages in between synchronise with LDS-only barriers.  The
// LDS that makes room for the copy comes from the last forward stage, whose twiddles are
// formed from the two-level table instead of read from a stage table.  Every LDS array
// is a separate static object, so the compiler's wait analysis keeps the copy apart from
// the transform's buffer.  grid (8 * per, C)
template <class P, bool BAND>
__global__ __launch_bounds__(P::nt, P::wpe_a) void k_a2_cols_pf(const float2* __restrict__ Z,
                                                              Optics o, fft::Plan plan,
                                                              const float2* __restrict__ tabs,
                                                              int per, float2* __restrict__ G) {
  constexpr int N = P::n, NT = P::nt, CG = P::cg;
  static_assert(CG == 1, "the prefetch path runs one column per workgroup");
  using Fw = typename P::Fwd;
  using Iv = typename P::Inv;
  using F1 = typename Fw::template G1<CG, NT>;
  using FL = typename Fw::template GL<CG, NT>;
  using IL = typename Iv::template GL<CG, NT>;
  constexpr int HALF = N / 2;                  // the crop holds at most N / 2 rows
  constexpr int TM = Fw::toff(Fw::L - 1);      // stage tables of the middle stages
  constexpr int NTW = (N + fft::TW_LO - 1) / fft::TW_LO + fft::TW_LO;
  __shared__ float2 buf[P::lds];
  __shared__ float2 stws[NTW];
  __shared__ float2 tt[TM];
  __shared__ float2 pfs[HALF];                 // the next plane's column, rows [0, oh)
  const int col0 = xcd_block(blockIdx.x, per);
  if (col0 >= o.nv) return;
  const int c = blockIdx.y;
  const float2* stw = stage_twiddles(plan, stws);
  {
    TableCopy<TM, NT> tcopy;
    tcopy.load(tabs);
    tcopy.store(tt);
  }
  // rows [0, oh) of plane d's column, stored as row pairs (z_at): lane l of wave w in
  // round p0 loads pair p = p0 + 64 w + l, rows 2p and 2p + 1, as one 16-byte piece into
  // pfs[2p], pfs[2p + 1] (global_load_lds_dwordx4: lane l writes bytes [16 l, 16 l + 16)
  // after the wave's LDS base)
  auto prefetch = [&](int d) {
    const float2* src = Z + z_at(o, d * o.C + c, 0, col0);
    const int np = o.oh >> 1;
    for (int p0 = 0; p0 < np; p0 += NT) {
      const int p = p0 + (int)threadIdx.x;
      const int wb = __builtin_amdgcn_readfirstlane(p0 + ((int)threadIdx.x & ~63));
      if (p < np)
        __builtin_amdgcn_global_load_lds(src + 2 * (size_t)((unsigned)(p * o.nv)),
                                         (__attribute__((address_space(3))) void*)(pfs + 2 * wb),
                                         16, 0, 0);
    }
  };
  float2 A[FL::MAXB][Fw::LAST];
  float Pg[FL::MAXB][Fw::LAST];
#pragma unroll
  for (int b = 0; b < FL::MAXB; ++b) {
    int j, col;
    FL::item(b, j, col);
#pragma unroll
    for (int k = 0; k < Fw::LAST; ++k) {
      A[b][k] = make_float2(0.f, 0.f);
      Pg[b][k] = FL::live(b) ? tf_point(o, c, j + k * FL::CNT, col0 + col) : -1.0f;
    }
  }
  prefetch(0);
  for (int d = 0; d < o.D; ++d) {
    __syncthreads();   // plane d is in pfs (the wait covers the copy); buf is free
    {
      float2 v[F1::MAXB][Fw::FIRST];
#pragma unroll
      for (int b = 0; b < F1::MAXB; ++b) {
        if (!F1::live(b)) continue;
        int j, col;
        F1::item(b, j, col);
#pragma unroll
        for (int k = 0; k < Fw::FIRST; ++k) {
          float2 u = make_float2(0.f, 0.f);
          if (in_mid_half<Fw>(k)) {
            const int rr = j + k * F1::CNT - o.or0;
            if ((unsigned)rr < (unsigned)o.oh) u = pfs[rr];
          }
          v[b][k] = u;
        }
      }
      Fw::template first<1, CG, NT>(buf, v);
    }
    fft::lds_barrier();   // every read of pfs is complete: it takes the next plane
    if (d + 1 < o.D) prefetch(d + 1);
    Fw::template middle<1, CG, NT, true>(buf, stw, plan.nh, tt);
    float2 v[FL::MAXB][Fw::LAST];
    Fw::template last<1, CG, NT, false>(buf, stw, plan.nh, v);
#pragma unroll
    for (int b = 0; b < FL::MAXB; ++b) {
      if (!FL::live(b)) continue;
      int j, col;
      FL::item(b, j, col);
#pragma unroll
      for (int k = 0; k < Fw::LAST; ++k) {
        const float2 h = tf_value<BAND>(o, Pg[b][k], c, d, j + k * FL::CNT, col0 + col);
        A[b][k].x += v[b][k].x * h.x + v[b][k].y * h.y;   // v * conj(h)
        A[b][k].y += v[b][k].y * h.x - v[b][k].x * h.y;
      }
    }
  }
  fft::lds_barrier();
  if (o.ap_tab) {
#pragma unroll
    for (int b = 0; b < FL::MAXB; ++b) {
      if (!FL::live(b)) continue;
      int j, col;
      FL::item(b, j, col);
#pragma unroll
      for (int k = 0; k < Fw::LAST; ++k)
        A[b][k] = fft::cscale(A[b][k], ap_square(o, j + k * FL::CNT, col0 + col));
    }
  }
  Iv::template first<-1, CG, NT>(buf, A);
  fft::lds_barrier();
  Iv::template middle<-1, CG, NT>(buf, stw, plan.nh);
  float2 u[IL::MAXB][Iv::LAST];
  Iv::template last<-1, CG, NT>(buf, stw, plan.nh, u);
  float2* dst = G + (size_t)c * o.H * o.nv + col0;
#pragma unroll
  for (int b = 0; b < IL::MAXB; ++b) {
    if (!IL::live(b)) continue;
    int j, col;
    IL::item(b, j, col);
#pragma unroll
    for (int k = Iv::LAST / 4; k < 3 * Iv::LAST / 4; ++k) {
      const int rr = j + k * IL::CNT - o.gr0;
      if ((unsigned)rr < (unsigned)o.H) dst[(unsigned)(rr * o.nv + col)] = u[b][k];
    }
  }
}

// ---------------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------------
// Row kernels on runtime plans, instantiated per radix class (<= 8, <= 17) and length
// class (<= 4608, <= 8704) so the per-thread butterfly registers match the transform.
#define ROW_LAUNCH(K, PLAN, GRID, SH, STREAM, ...)                                          \
  do {                                                                                      \
    const int n_ = (PLAN).len;                                                              \
    const bool small_ = n_ <= ROW_LCAP_SMALL;                                               \
    if (fft::max_radix(PLAN) <= 8) {                                                        \
      if (small_) launch_row(K<fft::Runtime<8, ROW_LCAP_SMALL>>, GRID, SH, STREAM, __VA_ARGS__); \
      else launch_row(K<fft::Runtime<8, ROW_LCAP>>, GRID, SH, STREAM, __VA_ARGS__);         \
    } else {                                                                                \
      if (small_) launch_row(K<fft::Runtime<17, ROW_LCAP_SMALL>>, GRID, SH, STREAM, __VA_ARGS__); \
      else launch_row(K<fft::Runtime<17, ROW_LCAP>>, GRID, SH, STREAM, __VA_ARGS__);        \
    }                                                                                       \
  } while (0)

// dynamic LDS above 64 KB must be enabled per kernel
template <class Kern>
static void allow_lds(Kern k, size_t bytes) {
  if (bytes > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <class Kern, class... Args>
static void launch_row(Kern k, dim3 grid, size_t sh, hipStream_t s, Args... args) {
  allow_lds(k, sh);
  hipLaunchKernelGGL(k, grid, dim3(ROW_NT), sh, s, args...);
}

static size_t row_lds_bytes(const fft::Plan& pc) {
  return (size_t)(fft::lds_elems(pc.len) + fft::twiddle_entries(pc.len)) * sizeof(float2);
}

// compile-time row plans of the standard workloads (4480: 1080p, 8320: 2160p, 1152: 256^2)
// 4480: layouts with at most 1/16 padding, so that four workgroups (39 KB of LDS each)
// fit on a CU
using RowP4480 = RowPlan<320, Pipe<4480, Rad<16, 8, 5, 7>, Lays<Ly<32, 1>, Ly<32, 2>, Ly0>>,
                         Pipe<4480, Rad<7, 5, 8, 16>, Lays<Ly0, Ly<64, 1>, Ly0>>>;
using RowP8320 = RowPlan<640, Pipe<8320, Rad<16, 8, 5, 13>, Lays<Ly<32, 1>, Ly<32, 4>, Ly0>>,
                         Pipe<8320, Rad<13, 5, 8, 16>, Lays<Ly0, Ly<32, 4>, Ly0>>>;
using RowP1152 = RowPlan<128, Pipe<1152, Rad<16, 8, 9>, Lays<Ly<32, 1>, Ly<32, 4>>>,
                         Pipe<1152, Rad<9, 8, 16>, Lays<Ly0, Ly<64, 1>>>>;
// row lengths of the batched configuration's frames (512^2, 1024x768, 1280x720)
using RowP1664 = RowPlan<128, Pipe<1664, Rad<16, 8, 13>, Lays<Ly<32, 1>, Ly<32, 2>>>,
                         Pipe<1664, Rad<13, 8, 16>, Lays<Ly0, Ly0>>>;
using RowP2688 = RowPlan<192, Pipe<2688, Rad<16, 8, 3, 7>, Lays<Ly<32, 1>, Ly<32, 2>, Ly0>>,
                         Pipe<2688, Rad<7, 3, 8, 16>, Lays<Ly0, Ly0, Ly0>>>;
using RowP3200 = RowPlan<256, Pipe<3200, Rad<16, 8, 5, 5>, Lays<Ly<32, 1>, Ly<32, 2>, Ly0>>,
                         Pipe<3200, Rad<5, 5, 8, 16>, Lays<Ly0, Ly0, Ly0>>>;
// compile-time column plans (2816: 1080p, 1152: 256^2, 4992: 2160p)
// 2816: the forward runs 2 columns per 512-thread workgroup (4 waves per SIMD, a small
// spill), the adjoint one column per 256-thread workgroup (3 waves per SIMD)
using ColP2816F = ColPlan<2, 512, CVQPG_WPE_F, 1, Pipe<2816, Rad<16, 16, 11>, Lays<Ly<32, 2>, Ly0>>,
                          Pipe<2816, Rad<11, 16, 16>, Lays<Ly0, Ly<32, 2>>>>;
using ColP2816A = ColPlan<1, 256, 1, CVQPG_WPE_A, Pipe<2816, Rad<16, 16, 11>, Lays<Ly<32, 1>, Ly<32, 2>>>,
                          Pipe<2816, Rad<11, 16, 16>, Lays<Ly0, Ly0>>>;
using ColP1152 = ColPlan<4, 512, 2, 2, Pipe<1152, Rad<16, 8, 9>, Lays<Ly<32, 2>, Ly0>>,
                         Pipe<1152, Rad<9, 8, 16>, Lays<Ly0, Ly<128, 2>>>>;
using ColP4992 = ColPlan<1, 384, 2, 2, Pipe<4992, Rad<16, 8, 3, 13>, Lays<Ly<32, 1>, Ly<32, 4>, Ly0>>,
                         Pipe<4992, Rad<13, 3, 8, 16>, Lays<Ly0, Ly0, Ly0>>>;
// column lengths of the batched configuration's frames (512^2, 1024x768, 1280x720)
using ColP1664 = ColPlan<2, 256, 1, 1, Pipe<1664, Rad<16, 8, 13>, Lays<Ly<32, 2>, Ly0>>,
                         Pipe<1664, Rad<13, 8, 16>, Lays<Ly0, Ly0>>>;
using ColP2176 = ColPlan<2, 320, 1, 1, Pipe<2176, Rad<16, 8, 17>, Lays<Ly<32, 2>, Ly0>>,
                         Pipe<2176, Rad<17, 8, 16>, Lays<Ly0, Ly0>>>;
using ColP2048 = ColPlan<2, 256, 1, 1, Pipe<2048, Rad<16, 16, 8>, Lays<Ly<32, 2>, Ly0>>,
                         Pipe<2048, Rad<8, 16, 16>, Lays<Ly<32, 2>, Ly<32, 2>>>>;

template <class P>
static size_t row_fx_lds_bytes() {
  return (size_t)(P::lds + fft::twiddle_entries(P::n)) * sizeof(float2);
}
template <class Fn>
static bool with_row_plan(int n, Fn&& fn) {
  switch (n) {
    case 4480: fn(RowP4480{}); return true;
    case 8320: fn(RowP8320{}); return true;
    case 1152: fn(RowP1152{}); return true;
    case 1664: fn(RowP1664{}); return true;
    case 2688: fn(RowP2688{}); return true;
    case 3200: fn(RowP3200{}); return true;
    default: return false;
  }
}

template <class P, class Kern, class... Args>
static void launch_fx(Kern k, dim3 grid, size_t sh, hipStream_t s, Args... args) {
  allow_lds(k, sh);
  hipLaunchKernelGGL(k, grid, dim3(P::nt), sh, s, args...);
}

template <class P>
static void launch_a3_fx(const Optics& o, const BlasmWS& w, const fft::Plan& pc,
                         const float* grad_intensity, const float2* of, hipStream_t s) {
  if (o.zpair) {
    constexpr size_t sh =
        (size_t)(a3_pair_lds<P>() + fft::twiddle_entries(P::n)) * sizeof(float2);
    allow_lds(k_a3_rows_pair_fx<P>, sh);
    hipLaunchKernelGGL(k_a3_rows_pair_fx<P>, dim3(o.oh / 2, o.D * o.C),
                       dim3(a3_pair_threads<P>()), sh, s, grad_intensity, of, o, pc, w.Y);
  } else {
    const size_t sh = (size_t)(P::lds + fft::twiddle_entries(P::n)) * sizeof(float2);
    allow_lds(k_a3_rows_fx<P>, sh);
    hipLaunchKernelGGL(k_a3_rows_fx<P>, dim3(o.oh, o.D * o.C), dim3(P::nt), sh, s,
                       grad_intensity, of, o, pc, w.Y);
  }
}

static int pick_cg(int nu, int nv) {
  for (int cg = 8; cg >= 1; cg >>= 1)
    if ((int64_t)nu * cg <= COL_LCAP && nv % cg == 0) return cg;
  return 0;
}

// Whether the adjoint column pass of a geometry runs the prefetching kernel
// (k_a2_cols_pf), i.e. launch_cols picks ColP2816A (column group 2, length 2816) or
// ColP4992 (column group 1, length 4992), both of one column per workgroup; that kernel
// reads the row spectra Z as row pairs, which the adjoint row pass then writes.
static bool adj_cols_prefetch(int nu, int cg) {
  return (cg == 2 && nu == 2816) || (cg == 1 && nu == 4992);
}

static int make_optics(Optics& o, int C, int H, int W, int pH, int pW, int D,
                       const double* wavelengths, const double* distances, double pitch,
                       double aperture_radius, int grid, int full_plane) {
  CVQPG_CHECK_ARG(C >= 1 && C <= 3, "channels must be 1..3 (got %d)", C);
  CVQPG_CHECK_ARG(D >= 1 && D <= MAXD, "planes must be 1..%d (got %d)", MAXD, D);
  CVQPG_CHECK_ARG(H > 0 && W > 0 && pH >= H && pW >= W, "invalid sizes");
  CVQPG_CHECK_ARG(pH % 2 == 0 && pW % 2 == 0 && H % 2 == 0 && W % 2 == 0,
                  "image and padded sizes must be even (odak centre crop)");
  CVQPG_CHECK_ARG(wavelengths && distances, "wavelengths/distances are NULL");
  CVQPG_CHECK_ARG(pitch > 0, "pixel pitch must be > 0");
  CVQPG_CHECK_ARG(grid == 0 || grid == 1, "grid must be 0 (odak) or 1 (fft)");
  o.C = C; o.D = D; o.H = H; o.W = W; o.pH = pH; o.pW = pW;
  o.nu = 2 * pH; o.nv = 2 * pW;
  const int mu = fft::stockham_len(o.nu), mv = fft::stockham_len(o.nv);
  CVQPG_CHECK_ARG(mv <= ROW_LCAP, "padded width %d exceeds the supported row length "
                  "(transform length %d, Stockham length %d > %d)", pW, o.nv, mv, ROW_LCAP);
  o.pad_h = (pH - H) / 2;
  o.pad_w = (pW - W) / 2;
  o.zr0 = pH - pH / 2;   // odak zero_pad: (2 pH)//2 - pH//2
  o.zc0 = pW - pW / 2;
  if (full_plane) {      // crop_center(2x -> padded): start n//4
    o.or0 = o.nu / 4; o.oc0 = o.nv / 4; o.oh = pH; o.ow = pW;
  } else {               // then crop_center(size = (H, W)) inside the padded plane
    o.or0 = o.nu / 4 + pH / 2 - H / 2; o.oc0 = o.nv / 4 + pW / 2 - W / 2; o.oh = H; o.ow = W;
  }
  o.gr0 = o.zr0 + o.pad_h;
  o.cg = pick_cg(mu, o.nv);
  CVQPG_CHECK_ARG(o.cg > 0, "padded height %d exceeds the supported column length "
                  "(transform length %d, Stockham length %d > %d)", pH, o.nu, mu, COL_LCAP);
  o.zpair = adj_cols_prefetch(o.nu, o.cg) ? 1 : 0;
  // the pruned transforms rely on every non-zero input and every kept output lying in
  // the centre half [n/4, 3n/4) of the 2x grid
  CVQPG_CHECK_ARG(o.zr0 == o.nu / 4 && o.zc0 == o.nv / 4 && o.or0 >= o.nu / 4 &&
                      o.or0 + o.oh <= 3 * (o.nu / 4) && o.oc0 >= o.nv / 4 &&
                      o.oc0 + o.ow <= 3 * (o.nv / 4),
                  "crop window outside the centre half of the padded grid");
  o.norm = (float)(1.0 / ((double)o.nu * (double)o.nv));
  const double X = pitch * o.nu, Y = pitch * o.nv;
  double fr0, frs, fc0, fcs;
  if (grid == 0) {  // odak: linspace(-1/(2dx) + 0.5/(2X), 1/(2dx) - 0.5/(2X), n)
    fr0 = -1.0 / (2 * pitch) + 0.5 / (2 * X);
    frs = (1.0 / pitch - 0.5 / X) / (o.nu - 1);
    fc0 = -1.0 / (2 * pitch) + 0.5 / (2 * Y);
    fcs = (1.0 / pitch - 0.5 / Y) / (o.nv - 1);
  } else {          // exact FFT bins (i - n/2) / X
    fr0 = -(double)(o.nu / 2) / X;
    frs = 1.0 / X;
    fc0 = -(double)(o.nv / 2) / Y;
    fcs = 1.0 / Y;
  }
  o.fr0 = (float)fr0; o.frs = (float)frs; o.fc0 = (float)fc0; o.fcs = (float)fcs;
  for (int c = 0; c < 3; ++c) {
    const double lam = c < C ? wavelengths[c] : 0.0;
    o.lam2[c] = (float)(lam * lam);
  }
  for (int d = 0; d < MAXD; ++d)
    for (int c = 0; c < 3; ++c) {
      if (d < D && c < C) {
        const double z = distances[d], lam = wavelengths[c];
        o.fxmax[d][c] = (float)(1.0 / sqrt((2 * z / X) * (2 * z / X) + 1) / lam);
        o.fymax[d][c] = (float)(1.0 / sqrt((2 * z / Y) * (2 * z / Y) + 1) / lam);
        const double zl = z / lam;
        o.zl[d][c] = (float)zl;
        const double fr = zl - floor(zl);
        o.carrier[d][c] = make_float2((float)cos(2 * M_PI * fr), (float)sin(2 * M_PI * fr));
      } else {
        o.fxmax[d][c] = o.fymax[d][c] = 0.0f;
        o.zl[d][c] = 0.0f;
        o.carrier[d][c] = make_float2(1.0f, 0.0f);
      }
    }
  {
    const double fymaxabs = fmax(fabs(fr0), fabs(fr0 + (o.nu - 1) * frs));
    const double fxmaxabs = fmax(fabs(fc0), fabs(fc0 + (o.nv - 1) * fcs));
    o.band_active = 0;
    for (int c = 0; c < C; ++c) {
      const double lam = wavelengths[c];
      const double x = lam * lam * (fxmaxabs * fxmaxabs + fymaxabs * fymaxabs);
      if (x >= 1.0) o.band_active = 1;
      // the largest residual phase on the grid, (z / lambda) g with g = x / (1 + sqrt(1 - x))
      // increasing in x: past 250 revolutions the kernels reduce it before the sine units
      const double gmax = x < 1.0 ? x / (1.0 + sqrt(1.0 - x)) : 1.0;
      for (int d = 0; d < D; ++d) {
        if (!((float)fxmaxabs < o.fxmax[d][c]) || !((float)fymaxabs < o.fymax[d][c]))
          o.band_active = 1;
        if (fabs(distances[d]) / lam * gmax > 250.0) o.band_active = 1;
      }
    }
  }
  // circular aperture on the centred grid: linspace(-n/2, n/2, n) per axis
  const double rmax = sqrt((o.nu / 2.0) * (o.nu / 2.0) + (o.nv / 2.0) * (o.nv / 2.0));
  o.ap_open = !(aperture_radius > 0) || aperture_radius > rmax + 1.0;
  o.ap_tab = nullptr;
  o.ap_r2 = (float)(aperture_radius * aperture_radius);
  o.apr0 = (float)(-o.nu / 2.0);
  o.aprs = (float)((double)o.nu / (o.nu - 1));
  o.apc0 = (float)(-o.nv / 2.0);
  o.apcs = (float)((double)o.nv / (o.nv - 1));
  return 0;
}

// Column kernels: compile-time plans for the column lengths of the standard workloads
// (with the column-group width pick_cg chooses for them), runtime plans otherwise.
template <int CG, class F>
static void launch_f2_k(const BlasmWS& w, const Optics& o, const fft::Plan& pr, hipStream_t s) {
  const int nb = o.nv / CG, per = ceil_div(nb, 8);
  const size_t sh = (size_t)(fft::lds_elems(pr.len * CG) + fft::twiddle_entries(pr.len)) * sizeof(float2);
  allow_lds(k_f2_cols<CG, F>, sh);
  hipLaunchKernelGGL((k_f2_cols<CG, F>), dim3(8 * per, o.C), dim3(COL_NT), sh, s, w.X, o, pr,
                     per, w.Y);
}

template <int CG, class F>
static void launch_a2_k(const BlasmWS& w, const Optics& o, const fft::Plan& pr, hipStream_t s) {
  const int nb = o.nv / CG, per = ceil_div(nb, 8);
  const size_t sh = (size_t)(fft::lds_elems(pr.len * CG) + fft::twiddle_entries(pr.len)) * sizeof(float2);
  allow_lds(k_a2_cols<CG, F>, sh);
  hipLaunchKernelGGL((k_a2_cols<CG, F>), dim3(8 * per, o.C), dim3(COL_NT), sh, s, w.Y, o, pr,
                     per, w.X);
}

template <int CG, bool ADJ, class F>
static void launch_cols_k(const BlasmWS& w, const Optics& o, const fft::Plan& pr, hipStream_t s) {
  if constexpr (ADJ) launch_a2_k<CG, F>(w, o, pr, s);
  else launch_f2_k<CG, F>(w, o, pr, s);
}

template <class P, bool ADJ>
static void launch_cols_fx(const BlasmWS& w, const Optics& o, const fft::Plan& pr, hipStream_t s,
                           int what) {
  static_assert(P::Fwd::TABLE + P::Inv::TABLE <= 8 * P::n, "stage tables exceed their region");
  constexpr int ne = P::Fwd::TABLE + P::Inv::TABLE;
  if (what & COLS_TABLES)
    hipLaunchKernelGGL(k_pipe_tables<P>, dim3(ceil_div(ne, 256)), dim3(256), 0, s, w.tw_cs);
  if (!(what & COLS_RUN)) return;
  const int nb = o.nv / P::cg, per = ceil_div(nb, 8);
  const int tab = ADJ ? P::Fwd::TABLE : P::Inv::TABLE;
  const size_t sh = (size_t)(col_lds_head<P>() + tab) * sizeof(float2);
  const dim3 grid(8 * per, o.C);
  const float2* tabs = w.tw_cs;
  if constexpr (ADJ && P::cg == 1) {   // all LDS of the prefetching kernel is static
    static_assert(std::is_same<P, ColP2816A>::value || std::is_same<P, ColP4992>::value,
                  "adj_cols_prefetch must name every plan of the prefetching kernel");
    if (o.band_active)
      launch_fx<P>(k_a2_cols_pf<P, true>, grid, 0, s, (const float2*)w.Y, o, pr, tabs, per, w.X);
    else
      launch_fx<P>(k_a2_cols_pf<P, false>, grid, 0, s, (const float2*)w.Y, o, pr, tabs, per, w.X);
  } else if constexpr (ADJ) {
    if (o.band_active)
      launch_fx<P>(k_a2_cols_fx<P, true>, grid, sh, s, (const float2*)w.Y, o, pr, tabs, per, w.X);
    else
      launch_fx<P>(k_a2_cols_fx<P, false>, grid, sh, s, (const float2*)w.Y, o, pr, tabs, per, w.X);
  } else {
    if (o.band_active)
      launch_fx<P>(k_f2_cols_fx<P, true>, grid, sh, s, (const float2*)w.X, o, pr, tabs, per, w.Y);
    else
      launch_fx<P>(k_f2_cols_fx<P, false>, grid, sh, s, (const float2*)w.X, o, pr, tabs, per, w.Y);
  }
}

template <int CG, bool ADJ>
static void launch_cols(const BlasmWS& w, const Optics& o, const fft::Plan& pr, hipStream_t s,
                        int what = COLS_TABLES | COLS_RUN) {
  // compile-time column plans of the standard workloads (2816: 1080p, 4992: 2160p, 1152: 256^2)
  if constexpr (CG == 2) {
    if (o.nu == 2816) {
      if constexpr (ADJ) return launch_cols_fx<ColP2816A, ADJ>(w, o, pr, s, what);
      else return launch_cols_fx<ColP2816F, ADJ>(w, o, pr, s, what);
    }
    if (o.nu == 1664) return launch_cols_fx<ColP1664, ADJ>(w, o, pr, s, what);
    if (o.nu == 2176) return launch_cols_fx<ColP2176, ADJ>(w, o, pr, s, what);
    if (o.nu == 2048) return launch_cols_fx<ColP2048, ADJ>(w, o, pr, s, what);
  }
  if constexpr (CG == 4) {
    if (o.nu == 1152) return launch_cols_fx<ColP1152, ADJ>(w, o, pr, s, what);
  }
  if constexpr (CG == 1) {
    if (o.nu == 4992) return launch_cols_fx<ColP4992, ADJ>(w, o, pr, s, what);
  }
  if (!(what & COLS_RUN)) return;   // runtime plans have no stage tables
  const int mr = fft::max_radix(pr);
  if (mr <= 8) launch_cols_k<CG, ADJ, fft::Runtime<8, COL_LCAP>>(w, o, pr, s);
  else if (mr <= 13) launch_cols_k<CG, ADJ, fft::Runtime<13, COL_LCAP>>(w, o, pr, s);
  else launch_cols_k<CG, ADJ, fft::Runtime<17, COL_LCAP>>(w, o, pr, s);
}

template <int CG>
static void launch_f2(const BlasmWS& w, const Optics& o, const fft::Plan& pr, hipStream_t s,
                      int what = COLS_TABLES | COLS_RUN) {
  launch_cols<CG, false>(w, o, pr, s, what);
}

template <int CG>
static void launch_a2(const BlasmWS& w, const Optics& o, const fft::Plan& pr, hipStream_t s,
                      int what = COLS_TABLES | COLS_RUN) {
  launch_cols<CG, true>(w, o, pr, s, what);
}

// The column kernels' stage tables depend only on the plan's radices, and the forward
// and the adjoint plan of a length share them, so one region serves both directions.
template <class A, class B>
constexpr bool same_pipe_radices() {
  if (A::L != B::L || A::TABLE != B::TABLE) return false;
  for (int i = 0; i < A::L; ++i)
    if (A::RR[i] != B::RR[i]) return false;
  return true;
}
static_assert(same_pipe_radices<ColP2816F::Fwd, ColP2816A::Fwd>() &&
                  same_pipe_radices<ColP2816F::Inv, ColP2816A::Inv>(),
              "the 2816-point forward and adjoint column plans must share stage tables");

static void launch_cols_for(const BlasmWS& w, const Optics& o, const fft::Plan& pr, hipStream_t s,
                            bool adj, int what) {
  switch (o.cg) {
    case 8: adj ? launch_a2<8>(w, o, pr, s, what) : launch_f2<8>(w, o, pr, s, what); break;
    case 4: adj ? launch_a2<4>(w, o, pr, s, what) : launch_f2<4>(w, o, pr, s, what); break;
    case 2: adj ? launch_a2<2>(w, o, pr, s, what) : launch_f2<2>(w, o, pr, s, what); break;
    default: adj ? launch_a2<1>(w, o, pr, s, what) : launch_f2<1>(w, o, pr, s, what); break;
  }
}

// Transform plans of both lengths; with `tables` also (re)computes their twiddle tables in
// the workspace (the first pass of a direction does; later passes reuse them).
static int plans(const Optics& o, const BlasmWS& w, fft::Plan& pr, fft::Plan& pc,
                 hipStream_t s, bool tables = true) {
  CVQPG_CHECK_ARG(fft::make_plan(o.nu, pr), "no FFT plan for length %d", o.nu);
  CVQPG_CHECK_ARG(fft::make_plan(o.nv, pc), "no FFT plan for length %d", o.nv);
  pr.tw = w.tw_r;
  pc.tw = w.tw_c;
  if (fft::is_bluestein(pr)) {
    pr.chirp = w.ch_r;
    pr.bhat = w.bh_r;
  }
  if (fft::is_bluestein(pc)) {
    pc.chirp = w.ch_c;
    pc.bhat = w.bh_c;
  }
  if (tables) {
    launch_cols_for(w, o, pr, s, false, COLS_TABLES);
    const int ne = std::max(fft::twiddle_entries(pr.len), fft::twiddle_entries(pc.len));
    hipLaunchKernelGGL(k_twiddles, dim3(ceil_div(ne, 256), 2), dim3(256), 0, s, w.tw_r, pr.len,
                       w.tw_c, pc.len);
    if (fft::is_bluestein(pr))
      hipLaunchKernelGGL(k_bluestein_tables, dim3(pr.len), dim3(256), 0, s, pr.n, pr.len, w.ch_r,
                         w.bh_r);
    if (fft::is_bluestein(pc))
      hipLaunchKernelGGL(k_bluestein_tables, dim3(pc.len), dim3(256), 0, s, pc.n, pc.len, w.ch_c,
                         w.bh_c);
  }
  return 0;
}

static int forward_cols_rows(const BlasmWS& w, const Optics& o, const fft::Plan& pr,
                             const fft::Plan& pc, float* out_field, float* intensity,
                             hipStream_t s, int passes = 6) {
  if (passes & 2) {
    launch_cols_for(w, o, pr, s, false, COLS_RUN);
    CVQPG_LAUNCH_CHECK();
  }
  if (!(passes & 4)) return 0;
  float2* of = reinterpret_cast<float2*>(out_field);
  const bool fx = with_row_plan(o.nv, [&](auto pl) {
    using P = decltype(pl);
    launch_fx<P>(k_f3_rows_fx<P>, dim3(o.oh, o.D * o.C), row_fx_lds_bytes<P>(), s, (const float2*)w.Y,
                 o, pc, of, intensity);
  });
  if (!fx) ROW_LAUNCH(k_f3_rows, pc, dim3(o.oh, o.D * o.C), row_lds_bytes(pc), s, w.Y, o, pc, of, intensity);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

}  // namespace cvqpg

using namespace cvqpg;

extern "C" size_t cvqpg_blasm_workspace_size(int channels, int planes, int height, int width,
                                             int pad_height, int pad_width, int full_plane) {
  const int oh = full_plane ? pad_height : height;
  (void)width;
  return carve_blasm(nullptr, channels, planes, height, pad_height, pad_width, oh).bytes;
}

extern "C" int cvqpg_blasm_forward_passes(int passes, const float* field, int channels,
                                          int height, int width, int pad_height, int pad_width,
                                          int planes, const double* wavelengths,
                                          const double* distances, double pixel_pitch,
                                          double aperture_radius, int grid, void* ws,
                                          size_t ws_bytes, float* out_field, float* intensity,
                                          void* stream) {
  CVQPG_CHECK_ARG(passes >= 0 && passes <= 31 && (passes & 7 || passes & CVQPG_BLASM_TABLES),
                  "passes must be a mask of 1, 2, 4, CVQPG_BLASM_TABLES and "
                  "CVQPG_BLASM_TABLES_KEPT (got %d)", passes);
  Optics o;
  if (int rc = make_optics(o, channels, height, width, pad_height, pad_width, planes,
                           wavelengths, distances, pixel_pitch, aperture_radius, grid, 0))
    return rc;
  CVQPG_CHECK_ARG((field || !(passes & 1)) && (intensity || !(passes & 4)) && ws,
                  "NULL field/intensity/workspace");
  const BlasmWS need = carve_blasm(nullptr, o.C, o.D, o.H, o.pH, o.pW, o.oh);
  CVQPG_CHECK_ARG(ws_bytes >= need.bytes, "workspace too small: %zu < %zu", ws_bytes, need.bytes);
  hipStream_t s = (hipStream_t)stream;
  const BlasmWS w = carve_blasm(ws, o.C, o.D, o.H, o.pH, o.pW, o.oh);
  fft::Plan pr, pc;
  if (int rc = plans(o, w, pr, pc, s, build_tables(passes))) return rc;
  if (passes & 1) {
    const float2* fl = reinterpret_cast<const float2*>(field);
    const bool fx = with_row_plan(o.nv, [&](auto pl) {
      using P = decltype(pl);
      launch_fx<P>(k_f1_rows_fx<P>, dim3(o.pH, o.C), row_fx_lds_bytes<P>(), s, fl, o, pc, w.X);
    });
    if (!fx) ROW_LAUNCH(k_f1_rows, pc, dim3(o.pH, o.C), row_lds_bytes(pc), s, fl, o, pc, w.X);
    CVQPG_LAUNCH_CHECK();
  }
  return forward_cols_rows(w, o, pr, pc, out_field, intensity, s, passes);
}

extern "C" int cvqpg_blasm_forward(const float* field, int channels, int height, int width,
                                   int pad_height, int pad_width, int planes,
                                   const double* wavelengths, const double* distances,
                                   double pixel_pitch, double aperture_radius, int grid,
                                   void* ws, size_t ws_bytes, float* out_field, float* intensity,
                                   void* stream) {
  return cvqpg_blasm_forward_passes(7, field, channels, height, width, pad_height, pad_width,
                                    planes, wavelengths, distances, pixel_pitch,
                                    aperture_radius, grid, ws, ws_bytes, out_field, intensity,
                                    stream);
}

extern "C" int cvqpg_blasm_backward_passes(int passes, const float* field,
                                           const float* grad_intensity, const float* out_field,
                                           int channels, int height, int width, int pad_height,
                                           int pad_width, int planes, const double* wavelengths,
                                           const double* distances, double pixel_pitch,
                                           double aperture_radius, int grid, void* ws,
                                           size_t ws_bytes, float* grad_field, void* stream) {
  CVQPG_CHECK_ARG(passes >= 0 && passes <= 31 && (passes & 7 || passes & CVQPG_BLASM_TABLES),
                  "passes must be a mask of 1, 2, 4, CVQPG_BLASM_TABLES and "
                  "CVQPG_BLASM_TABLES_KEPT (got %d)", passes);
  Optics o;
  if (int rc = make_optics(o, channels, height, width, pad_height, pad_width, planes,
                           wavelengths, distances, pixel_pitch, aperture_radius, grid, 0))
    return rc;
  CVQPG_CHECK_ARG(((field && grad_intensity && out_field && grad_field) || !(passes & 7)) && ws,
                  "NULL argument pointer");
  const BlasmWS need = carve_blasm(nullptr, o.C, o.D, o.H, o.pH, o.pW, o.oh);
  CVQPG_CHECK_ARG(ws_bytes >= need.bytes, "workspace too small: %zu < %zu", ws_bytes, need.bytes);
  hipStream_t s = (hipStream_t)stream;
  const BlasmWS w = carve_blasm(ws, o.C, o.D, o.H, o.pH, o.pW, o.oh);
  fft::Plan pr, pc;
  if (int rc = plans(o, w, pr, pc, s, build_tables(passes))) return rc;
  if (passes & 1) {
    const float2* of = reinterpret_cast<const float2*>(out_field);
    const bool fx3 = with_row_plan(o.nv, [&](auto pl) {
      using P = decltype(pl);
      launch_a3_fx<P>(o, w, pc, grad_intensity, of, s);
    });
    if (!fx3) ROW_LAUNCH(k_a3_rows, pc, dim3(o.oh, o.D * o.C), row_lds_bytes(pc), s, grad_intensity, of, o, pc, w.Y);
    CVQPG_LAUNCH_CHECK();
  }
  if (passes & 2) {
    launch_cols_for(w, o, pr, s, true, COLS_RUN);
    CVQPG_LAUNCH_CHECK();
  }
  if (passes & 4) {
    const float2* fl = reinterpret_cast<const float2*>(field);
    float2* gf = reinterpret_cast<float2*>(grad_field);
    const bool fx1 = with_row_plan(o.nv, [&](auto pl) {
      using P = decltype(pl);
      launch_fx<P>(k_a1_rows_fx<P>, dim3(o.H, o.C), row_fx_lds_bytes<P>(), s, (const float2*)w.X,
                   fl, o, pc, gf);
    });
    if (!fx1) ROW_LAUNCH(k_a1_rows, pc, dim3(o.H, o.C), row_lds_bytes(pc), s, w.X, fl, o, pc, gf);
    CVQPG_LAUNCH_CHECK();
  }
  return 0;
}

extern "C" int cvqpg_blasm_backward(const float* field, const float* grad_intensity,
                                    const float* out_field, int channels, int height, int width,
                                    int pad_height, int pad_width, int planes,
                                    const double* wavelengths, const double* distances,
                                    double pixel_pitch, double aperture_radius, int grid,
                                    void* ws, size_t ws_bytes, float* grad_field, void* stream) {
  return cvqpg_blasm_backward_passes(7, field, grad_intensity, out_field, channels, height, width,
                                     pad_height, pad_width, planes, wavelengths, distances,
                                     pixel_pitch, aperture_radius, grid, ws, ws_bytes, grad_field,
                                     stream);
}

static int reconstruct_impl(const float* phase, const float* amplitude, const float* aperture,
                            int channels, int pad_height, int pad_width, int planes,
                            const double* wavelengths, const double* distances,
                            double pixel_pitch, double aperture_radius, int grid, void* ws,
                            size_t ws_bytes, float* out_field, float* intensity, void* stream) {
  Optics o;
  if (int rc = make_optics(o, channels, pad_height, pad_width, pad_height, pad_width, planes,
                           wavelengths, distances, pixel_pitch, aperture_radius, grid, 1))
    return rc;
  o.ap_tab = aperture;
  CVQPG_CHECK_ARG(phase && amplitude && intensity && ws, "NULL argument pointer");
  const BlasmWS need = carve_blasm(nullptr, o.C, o.D, o.H, o.pH, o.pW, o.oh);
  CVQPG_CHECK_ARG(ws_bytes >= need.bytes, "workspace too small: %zu < %zu", ws_bytes, need.bytes);
  hipStream_t s = (hipStream_t)stream;
  const BlasmWS w = carve_blasm(ws, o.C, o.D, o.H, o.pH, o.pW, o.oh);
  fft::Plan pr, pc;
  if (int rc = plans(o, w, pr, pc, s)) return rc;
  ROW_LAUNCH(k_f1_rows_polar, pc, dim3(o.pH, o.C), row_lds_bytes(pc), s, phase, amplitude, o, pc, w.X);
  CVQPG_LAUNCH_CHECK();
  return forward_cols_rows(w, o, pr, pc, out_field, intensity, s);
}

static int reconstruct_backward_impl(
    const float* phase, const float* amplitude, const float* aperture,
    const float* grad_intensity, const float* out_field, int channels, int pad_height,
    int pad_width, int planes, const double* wavelengths, const double* distances,
    double pixel_pitch, double aperture_radius, int grid, void* ws, size_t ws_bytes,
    float* grad_phase, float* grad_amplitude, void* stream) {
  Optics o;
  if (int rc = make_optics(o, channels, pad_height, pad_width, pad_height, pad_width, planes,
                           wavelengths, distances, pixel_pitch, aperture_radius, grid, 1))
    return rc;
  o.ap_tab = aperture;
  CVQPG_CHECK_ARG(phase && amplitude && grad_intensity && out_field && grad_phase &&
                      grad_amplitude && ws,
                  "NULL argument pointer");
  const BlasmWS need = carve_blasm(nullptr, o.C, o.D, o.H, o.pH, o.pW, o.oh);
  CVQPG_CHECK_ARG(ws_bytes >= need.bytes, "workspace too small: %zu < %zu", ws_bytes, need.bytes);
  hipStream_t s = (hipStream_t)stream;
  const BlasmWS w = carve_blasm(ws, o.C, o.D, o.H, o.pH, o.pW, o.oh);
  fft::Plan pr, pc;
  if (int rc = plans(o, w, pr, pc, s)) return rc;
  const float2* of = reinterpret_cast<const float2*>(out_field);
  const bool fx3 = with_row_plan(o.nv, [&](auto pl) {
    using P = decltype(pl);
    launch_a3_fx<P>(o, w, pc, grad_intensity, of, s);
  });
  if (!fx3) ROW_LAUNCH(k_a3_rows, pc, dim3(o.oh, o.D * o.C), row_lds_bytes(pc), s, grad_intensity, of, o, pc, w.Y);
  launch_cols_for(w, o, pr, s, true, COLS_RUN);
  ROW_LAUNCH(k_a1_rows_polar, pc, dim3(o.pH, o.C), row_lds_bytes(pc), s, (const float2*)w.X, phase,
             amplitude, o, pc, grad_phase, grad_amplitude);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_blasm_reconstruct(const float* phase, const float* amplitude, int channels,
                                       int pad_height, int pad_width, int planes,
                                       const double* wavelengths, const double* distances,
                                       double pixel_pitch, double aperture_radius, int grid,
                                       void* ws, size_t ws_bytes, float* out_field,
                                       float* intensity, void* stream) {
  return reconstruct_impl(phase, amplitude, nullptr, channels, pad_height, pad_width, planes,
                          wavelengths, distances, pixel_pitch, aperture_radius, grid, ws,
                          ws_bytes, out_field, intensity, stream);
}

extern "C" int cvqpg_blasm_reconstruct_backward(
    const float* phase, const float* amplitude, const float* grad_intensity,
    const float* out_field, int channels, int pad_height, int pad_width, int planes,
    const double* wavelengths, const double* distances, double pixel_pitch,
    double aperture_radius, int grid, void* ws, size_t ws_bytes, float* grad_phase,
    float* grad_amplitude, void* stream) {
  return reconstruct_backward_impl(phase, amplitude, nullptr, grad_intensity, out_field,
                                   channels, pad_height, pad_width, planes, wavelengths,
                                   distances, pixel_pitch, aperture_radius, grid, ws, ws_bytes,
                                   grad_phase, grad_amplitude, stream);
}

extern "C" int cvqpg_blasm_reconstruct_aperture(
    const float* phase, const float* amplitude, const float* aperture, int channels,
    int pad_height, int pad_width, int planes, const double* wavelengths,
    const double* distances, double pixel_pitch, int grid, void* ws, size_t ws_bytes,
    float* out_field, float* intensity, void* stream) {
  CVQPG_CHECK_ARG(aperture, "NULL aperture");
  return reconstruct_impl(phase, amplitude, aperture, channels, pad_height, pad_width, planes,
                          wavelengths, distances, pixel_pitch, 0.0, grid, ws, ws_bytes,
                          out_field, intensity, stream);
}

extern "C" int cvqpg_blasm_reconstruct_aperture_backward(
    const float* phase, const float* amplitude, const float* aperture,
    const float* grad_intensity, const float* out_field, int channels, int pad_height,
    int pad_width, int planes, const double* wavelengths, const double* distances,
    double pixel_pitch, int grid, void* ws, size_t ws_bytes, float* grad_phase,
    float* grad_amplitude, void* stream) {
  CVQPG_CHECK_ARG(aperture, "NULL aperture");
  return reconstruct_backward_impl(phase, amplitude, aperture, grad_intensity, out_field,
                                   channels, pad_height, pad_width, planes, wavelengths,
                                   distances, pixel_pitch, 0.0, grid, ws, ws_bytes, grad_phase,
                                   grad_amplitude, stream);
}
