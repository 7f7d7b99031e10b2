// Tile binning, stable LDS radix sort and the quadratic-phase Gaussian raster
// (forward and adjoint) for gfx950.
//
// Reference behaviour (file:line in the reference repository):
//   rect / count        cuda/forward.cu:8-26
//   key emission        cuda/forward.cu:28-47
//   sort + ranges       cuda/raster.cu:77-95, cuda/forward.cu:49-60
//   forward render      cuda/forward.cu:62-162
//   adjoint             cuda/backward.cu:22-168
//
// Design (MI355X-first, not a translation):
//   * keys are 32-bit tile ids with the Gaussian id as the payload; emission is in
//     ascending id order, so a stable radix sort on ceil(log2 T) bits reproduces
//     the reference's 64-bit (tile << 32 | id) order exactly;
//   * the sort is a 3-kernel LSD radix sort per 8-bit digit (LDS histogram,
//     device scan, LDS-ranked stable scatter using 64-lane ballots);
//   * forward: one 256-thread workgroup per 16x16 tile, one pixel per lane,
//     Gaussians staged through LDS in batches of 256 and read as wave-uniform
//     (broadcast) ds_read_b128; base-2 exp and revolution-scaled sin/cos
//     (v_exp_f32, v_sin_f32, v_cos_f32); complex64 stored once per pixel,
//     no atomics;
//   * adjoint: the transpose of the forward's parallelisation -- one Gaussian per
//     lane, the tile's pixel gradients staged in LDS and read as broadcasts, so
//     every lane accumulates its own Gaussian's partials in registers with no
//     cross-lane reduction; one 64-byte partial record per (tile, Gaussian) key
//     is written at the key's emission slot and a per-Gaussian pass sums the
//     contiguous slots (deterministic, no atomics).
#include <string.h>

#include <type_traits>

#include "common.h"
#include "cvqpg.h"

namespace cvqpg {

constexpr int TILE = 16;
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_BLOCK = SCAN_THREADS * SCAN_ITEMS;  // 1024 elements per block
constexpr int SORT_THREADS = 256;
constexpr int SORT_ROUNDS = 16;
constexpr int SORT_BLOCK = SORT_THREADS * SORT_ROUNDS;  // 4096 keys per block
constexpr int REC = 16;                                 // floats per packed Gaussian / partial record
// G = exp(-mahal/2) is evaluated as 2^q with q = -mahal/2 * log2(e).  The reference clamps
// the exponent at -50 (q = -72.1); pairs there contribute below e^-50 ~ 2e-22 and the
// kernels evaluate them as 2^q unclamped (see the comments at the uses).
constexpr float LOG2E = 1.4426950408889634f;
constexpr float QSCALE = -0.5f * LOG2E;
// forward wave skip: all pixels of the wave below G = 2^-24
constexpr float QSKIP = -24.0f;
constexpr float INV_2PI = 0.15915494309189535f;

__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fract_f(float x) { return __builtin_amdgcn_fractf(x); }
__device__ __forceinline__ float sin_rev(float x) { return __builtin_amdgcn_sinf(x); }
__device__ __forceinline__ float cos_rev(float x) { return __builtin_amdgcn_cosf(x); }

// The sine and cosine units take revolutions over [-256, 256]: measured on 8.2 M inputs
// over that range, their results lie within 6e-8 of those for the input's fractional part
// (the phase u + w q itself is rounded to half a unit in the last place of |u + w q|
// revolutions, up to 7.6e-6 near 250), and on 0.4 M inputs up to 1e9 in magnitude past
// it they stay finite and within [-1, 1], unrelated to the angle.  The terms that count
// have q >= QSKIP, so when |u| + |QSKIP| |w| <= REV_DIRECT for every channel of a
// Gaussian its phase goes to the units unreduced: terms with q < QSKIP may then take a
// wrong angle, and stay below 2^-24 of a unit term like the terms the kernels leave out.
// Other Gaussians keep the fract.
constexpr float REV_DIRECT = 250.0f;
__device__ __forceinline__ bool rev_direct(float u, float w) {
  return fabsf(u) + (-QSKIP) * fabsf(w) <= REV_DIRECT;
}
// FR = false: the phase of a Gaussian that allows it (all of a batch or wave do) goes to the
// units unreduced; FR = true: each Gaussian keeps its own choice (direct), so a term's bits
// depend only on its own Gaussian, not on the batch or wave it was evaluated in.
template <bool FR>
__device__ __forceinline__ float rev_of(float x, bool direct) {
  return FR && !direct ? fract_f(x) : x;
}

// ------------------------------------------------------------------------------
// workspace
// ------------------------------------------------------------------------------
struct RasterWS {
  // regions that depend on n only come first, so a workspace carved for a larger
  // capacity keeps the same binning-count layout
  float* packed;        // [n][REC]   mx, my, a00, a01, a11, o, beta, (u_c, w_c, col_c) x C
  int4* rect;           // [n]        (emission offset, lo_x, lo_y, rect width)
  uint32_t* counts;     // [n]        tiles touched
  uint32_t* offs;       // [n]        exclusive prefix of counts
  uint32_t* sums_n;     // scan block sums of the count scan
  uint64_t* live;       // [n]        tiles of the rectangle a live count keeps (bit per tile)
  uint32_t* misc;       // [0] total pairs, [1] overflow flag of the last count,
                        // [2] overflowed counts since the last clear (sticky),
                        // [3] 1 if the last count was cvqpg_bin_count_live
  // capacity-dependent regions
  uint32_t* keys_a;     // [cap]
  uint32_t* vals_a;     // [cap]
  uint32_t* keys_b;     // [cap]
  uint32_t* vals_b;     // [cap]
  uint32_t* hist;       // [256 * nb]
  uint32_t* sums_h;     // scan block sums of the histogram scans
  float* partials;      // [cap][REC]
  size_t bytes;
};

static int sort_blocks(int64_t cap) { return ceil_div(cap > 0 ? cap : 1, SORT_BLOCK); }

static RasterWS carve(void* base, int n, int64_t cap) {
  Carver cv(base);
  RasterWS w;
  int nn = n > 0 ? n : 1;
  int64_t cc = cap > 0 ? cap : 1;
  int nb = sort_blocks(cc);
  w.packed = cv.take<float>((size_t)nn * REC);
  w.rect = cv.take<int4>(nn);
  w.counts = cv.take<uint32_t>(nn);
  w.offs = cv.take<uint32_t>(nn);
  w.sums_n = cv.take<uint32_t>(ceil_div(nn, SCAN_BLOCK) + 1);
  w.live = cv.take<uint64_t>(nn);
  w.misc = cv.take<uint32_t>(64);
  w.keys_a = cv.take<uint32_t>(cc);
  w.vals_a = cv.take<uint32_t>(cc);
  w.keys_b = cv.take<uint32_t>(cc);
  w.vals_b = cv.take<uint32_t>(cc);
  w.hist = cv.take<uint32_t>((size_t)256 * nb);
  w.sums_h = cv.take<uint32_t>(ceil_div((int64_t)256 * nb, SCAN_BLOCK) + 1);
  w.partials = cv.take<float>((size_t)cc * REC);
  w.bytes = cv.bytes();
  return w;
}

// ------------------------------------------------------------------------------
// generic device-wide exclusive scan of uint32 (3 kernels)
// ------------------------------------------------------------------------------
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds, int nthreads, uint32_t& total) {
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int off = 1; off < nthreads; off <<= 1) {
    uint32_t x = (t >= off) ? lds[t - off] : 0u;
    __syncthreads();
    lds[t] += x;
    __syncthreads();
  }
  const uint32_t incl = lds[t];
  total = lds[nthreads - 1];
  __syncthreads();
  return incl - v;
}

// per block of 4096: out = local exclusive prefix, sums[b] = block total
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_local(const uint32_t* __restrict__ in,
                                                             uint32_t* __restrict__ out,
                                                             uint32_t* __restrict__ sums,
                                                             int64_t n) {
  __shared__ uint32_t tile[SCAN_BLOCK + SCAN_BLOCK / 16];
  __shared__ uint32_t red[SCAN_THREADS];
  const int t = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    const int li = i * SCAN_THREADS + t;
    const int64_t gi = base + li;
    tile[li + (li >> 4)] = gi < n ? in[gi] : 0u;
  }
  __syncthreads();
  uint32_t vals[SCAN_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    const int li = t * SCAN_ITEMS + i;
    vals[i] = tile[li + (li >> 4)];
    s += vals[i];
  }
  uint32_t total;
  uint32_t run = block_excl_scan(s, red, SCAN_THREADS, total);
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    const int li = t * SCAN_ITEMS + i;
    tile[li + (li >> 4)] = run;
    run += vals[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    const int li = i * SCAN_THREADS + t;
    const int64_t gi = base + li;
    if (gi < n) out[gi] = tile[li + (li >> 4)];
  }
  if (t == 0) sums[blockIdx.x] = total;
}

// single block: exclusive scan of the block sums; optional total / overflow flag
__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t* __restrict__ sums, int nb,
                                                    uint32_t* __restrict__ misc, int64_t cap,
                                                    int32_t* __restrict__ num_pairs,
                                                    int32_t* __restrict__ point_list0) {
  __shared__ uint32_t red[1024];
  const int t = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int b0 = t * per;
  uint32_t s = 0;
  for (int i = 0; i < per; ++i)
    if (b0 + i < nb) s += sums[b0 + i];
  uint32_t total;
  uint32_t run = block_excl_scan(s, red, 1024, total);
  for (int i = 0; i < per; ++i)
    if (b0 + i < nb) {
      const uint32_t v = sums[b0 + i];
      sums[b0 + i] = run;
      run += v;
    }
  if (t == 0 && misc) {
    const uint32_t over = (int64_t)total > cap ? 1u : 0u;
    misc[0] = total;
    misc[1] = over;
    misc[2] += over;
    if (num_pairs) *num_pairs = (int32_t)total;
    if (point_list0 && total == 0) *point_list0 = 0;
  }
}

__global__ void k_scan_add(uint32_t* __restrict__ out, const uint32_t* __restrict__ sums,
                           int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += sums[i / SCAN_BLOCK];
}

// Second half of the exclusive scan when there are at most 1024 block sums: every
// workgroup adds to its 256 elements the sum of the block sums before their SCAN_BLOCK
// chunk, reducing them itself (so no separate pass scans the block sums); the first
// workgroup, whose prefix is 0, reduces all of them for the total, the overflow flag, the
// key count and the empty list's point_list[0] = 0.  grid ceil(n / 256)
__global__ __launch_bounds__(256) void k_scan_finish(uint32_t* __restrict__ out,
                                                     const uint32_t* __restrict__ sums, int nb,
                                                     int64_t n, uint32_t* __restrict__ misc,
                                                     int64_t cap, int32_t* __restrict__ num_pairs,
                                                     int32_t* __restrict__ point_list0) {
  static_assert(SCAN_BLOCK % 256 == 0, "a workgroup's elements lie in one chunk");
  __shared__ uint32_t red[256];
  const int t = threadIdx.x;
  const int64_t first = (int64_t)blockIdx.x * 256;
  const int chunk = (int)(first / SCAN_BLOCK);
  const int lim = blockIdx.x == 0 ? nb : chunk;
  uint32_t s = 0;
  for (int k = t; k < lim; k += 256) s += sums[k];
  red[t] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  const uint32_t pre = red[0];
  const int64_t i = first + t;
  if (blockIdx.x == 0) {
    if (t == 0 && misc) {
      const uint32_t over = (int64_t)pre > cap ? 1u : 0u;
      misc[0] = pre;
      misc[1] = over;
      misc[2] += over;
      if (num_pairs) *num_pairs = (int32_t)pre;
      if (point_list0 && pre == 0) *point_list0 = 0;
    }
  } else if (i < n) {
    out[i] += pre;
  }
}

static int scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t* sums, int64_t n,
                          uint32_t* misc, int64_t cap, int32_t* num_pairs, int32_t* pl0,
                          hipStream_t s) {
  const int nb = ceil_div(n > 0 ? n : 1, SCAN_BLOCK);
  CVQPG_CHECK_ARG(nb <= 1024 * 64, "scan of %lld elements exceeds the supported size",
                  (long long)n);
  hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(SCAN_THREADS), 0, s, in, out, sums, n);
  if (nb <= 1024) {
    hipLaunchKernelGGL(k_scan_finish, dim3(ceil_div(n > 0 ? n : 1, 256)), dim3(256), 0, s, out,
                       sums, nb, n, misc, cap, num_pairs, pl0);
  } else {
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, sums, nb, misc, cap, num_pairs,
                       pl0);
    hipLaunchKernelGGL(k_scan_add, dim3(ceil_div(n > 0 ? n : 1, 256)), dim3(256), 0, s, out,
                       sums, n);
  }
  CVQPG_LAUNCH_CHECK();
  return 0;
}

// ------------------------------------------------------------------------------
// binning
// ------------------------------------------------------------------------------
// cuda/forward.cu:8-13, evaluated with the same float32 operation sequence.
__device__ __forceinline__ void get_rect(float px, float py, int r, int gx, int gy, int& lx,
                                         int& ly, int& hx, int& hy) {
  const float rf = (float)r;
  const float b = (float)TILE;
  lx = min(gx, max(0, (int)(__fsub_rn(px, rf) / b)));
  ly = min(gy, max(0, (int)(__fsub_rn(py, rf) / b)));
  hx = min(gx, max(0, (int)(__fsub_rn(__fadd_rn(__fadd_rn(px, rf), b), 1.0f) / b)));
  hy = min(gy, max(0, (int)(__fsub_rn(__fadd_rn(__fadd_rn(py, rf), b), 1.0f) / b)));
}

__global__ void k_count(const float2* __restrict__ means, const int32_t* __restrict__ radii,
                        int n, int gx, int gy, uint32_t* __restrict__ counts,
                        int4* __restrict__ rect, uint32_t* __restrict__ misc) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0) misc[3] = 0u;  // every tile of the rectangles (the reference's lists)
  if (g >= n) return;
  const int r = radii[g];
  uint32_t c = 0;
  int lx = 0, ly = 0, hx = 0, hy = 0;
  if (r > 0) {
    const float2 m = means[g];
    get_rect(m.x, m.y, r, gx, gy, lx, ly, hx, hy);
    if (hx > lx && hy > ly) c = (uint32_t)((hy - ly) * (hx - lx));
  }
  counts[g] = c;
  rect[g] = make_int4(0, lx, ly, c ? hx - lx : 0);
}

// Position of the k-th set bit (from 0) of m, which has more than k set bits.
__device__ __forceinline__ uint32_t select_bit(uint64_t m, uint32_t k) {
  uint32_t pos = 0;
#pragma unroll
  for (int width = 32; width >= 1; width >>= 1) {
    const uint32_t c = (uint32_t)__popcll(m & ((1ull << width) - 1ull));
    if (k >= c) {
      k -= c;
      m >>= width;
      pos += (uint32_t)width;
    }
  }
  return pos;
}

// The keys of one workgroup's 256 Gaussians fill one contiguous range of the key array
// (their exclusive offsets are consecutive), so the workgroup emits them together:
// thread t writes entries t, t + 256, ... of the range (coalesced), finding the entry's
// Gaussian by a binary search over the workgroup's offsets in LDS.  Each Gaussian's keys
// run over its rectangle row by row, as in the reference.  After a live count (misc[3])
// a Gaussian's keys are the kept tiles of its rectangle (the set bits of live[g], in the
// same row-major order; rectangles of more than 64 tiles keep them all).  The grid first
// clears the tile ranges (untouched tiles keep (0, 0)) and sets point_list[0] = 0 (the
// reference's list when no key exists).  grid ceil(n / 256)
__global__ __launch_bounds__(256) void k_duplicate(const uint32_t* __restrict__ counts,
                                                   const uint32_t* __restrict__ offs,
                                                   int4* __restrict__ rect,
                                                   const uint64_t* __restrict__ live,
                                                   const uint32_t* __restrict__ misc, int n,
                                                   int gx, uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals,
                                                   int32_t* __restrict__ ranges, int64_t nr,
                                                   int32_t* __restrict__ point_list) {
  __shared__ uint32_t s_off[257];  // offsets relative to the range start; [256] = its end
  __shared__ int4 s_rc[256];
  __shared__ uint64_t s_live[256];
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nr; i += (int64_t)gridDim.x * 256)
    ranges[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) point_list[0] = 0;
  if (misc[1]) return;  // uniform: the keys overflowed the capacity
  const bool lv = misc[3] != 0u;
  const int t = threadIdx.x, g0 = blockIdx.x * 256, g = g0 + t;
  const int last = min(g0 + 255, n - 1);
  const uint32_t base = offs[g0];
  const uint32_t end = offs[last] + counts[last] - base;
  int4 rc = make_int4(0, 0, 0, 1);
  uint32_t rel = end;
  if (g < n) {
    rel = offs[g] - base;
    rc = rect[g];
    rc.x = (int)(base + rel);
    rect[g] = rc;
  }
  s_off[t] = rel;
  s_rc[t] = rc;
  s_live[t] = (lv && g < n) ? live[g] : 0ull;
  if (t == 0) s_off[256] = end;
  __syncthreads();
  for (uint32_t i = t; i < end; i += 256) {
    int lo = 0, hi = 255;  // the last j with s_off[j] <= i has s_off[j + 1] > i
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_off[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    const int4 r = s_rc[lo];
    const uint32_t local = i - s_off[lo];
    const uint32_t pos =
        (lv && s_off[lo + 1] - s_off[lo] <= 64u) ? select_bit(s_live[lo], local) : local;
    const uint32_t y = pos / (uint32_t)r.w, x = pos - y * (uint32_t)r.w;
    keys[base + i] = (uint32_t)(r.z + (int)y) * (uint32_t)gx + (uint32_t)r.y + x;
    vals[base + i] = (uint32_t)(g0 + lo);
  }
}

// ------------------------------------------------------------------------------
// stable LSD radix sort on the tile id (8-bit digits)
// ------------------------------------------------------------------------------
__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                             const uint32_t* __restrict__ misc,
                                                             int shift, uint32_t mask, int nb,
                                                             uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  const int t = threadIdx.x;
  h[t] = 0;
  __syncthreads();
  const int64_t P = misc[1] ? 0 : (int64_t)misc[0];
  const int64_t base = (int64_t)blockIdx.x * SORT_BLOCK;
  // all the block's keys are loaded before any is counted (a guarded load and its use
  // would wait for each key in turn); keys past the end read element 0 and are not counted
  uint32_t kv[SORT_ROUNDS];
#pragma unroll
  for (int i = 0; i < SORT_ROUNDS; ++i) {
    const int64_t idx = base + i * SORT_THREADS + t;
    kv[i] = keys[idx < P ? idx : 0];
  }
#pragma unroll
  for (int i = 0; i < SORT_ROUNDS; ++i) {
    const int64_t idx = base + i * SORT_THREADS + t;
    if (idx < P) atomicAdd(&h[(kv[i] >> shift) & mask], 1u);
  }
  __syncthreads();
  hist[(int64_t)t * nb + blockIdx.x] = h[t];
}

__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ misc, int shift, int nbits,
    int nb) {
  __shared__ uint32_t digit_base[256];
  __shared__ uint32_t wcnt[4][256];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int64_t P = misc[1] ? 0 : (int64_t)misc[0];
  const int64_t blk = (int64_t)blockIdx.x * SORT_BLOCK;
  if (blk >= P) return;  // uniform over the block
  digit_base[t] = hist[(int64_t)t * nb + blockIdx.x];
#pragma unroll
  for (int w = 0; w < 4; ++w) wcnt[w][t] = 0;
  const uint32_t mask = (1u << nbits) - 1u;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // the block's keys and values are all loaded up front (one memory latency instead of
  // one per round); entries past the end read element 0 and are not placed
  uint32_t kk[SORT_ROUNDS], vv[SORT_ROUNDS];
#pragma unroll
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const int64_t idx = blk + r * SORT_THREADS + t;
    const int64_t ii = idx < P ? idx : 0;
    kk[r] = keys_in[ii];
    vv[r] = vals_in[ii];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const int64_t idx = blk + r * SORT_THREADS + t;
    const bool valid = idx < P;
    const uint32_t k = valid ? kk[r] : 0u;
    const uint32_t v = valid ? vv[r] : 0u;
    const uint32_t d = (k >> shift) & mask;
    uint64_t m = __ballot(valid);
    for (int b = 0; b < nbits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(m & lt);
    if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(m);
    __syncthreads();
    if (valid) {
      uint32_t pos = digit_base[d] + rank;
      for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
      keys_out[pos] = k;
      vals_out[pos] = v;
    }
    __syncthreads();
    digit_base[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
#pragma unroll
    for (int w = 0; w < 4; ++w) wcnt[w][t] = 0;
    __syncthreads();
  }
}

__global__ void k_clear_overflow(uint32_t* __restrict__ misc) {
  if (threadIdx.x < 3) misc[threadIdx.x] = 0u;
}

__global__ void k_clear_ranges(int32_t* __restrict__ ranges, int64_t n,
                               int32_t* __restrict__ point_list) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ranges[i] = 0;
  if (i == 0) point_list[0] = 0;
}

__global__ void k_copy_u32(const uint32_t* __restrict__ src, int32_t* __restrict__ dst,
                           const uint32_t* __restrict__ misc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t P = misc[1] ? 0 : (int64_t)misc[0];
  if (i < P) dst[i] = (int32_t)src[i];
}

__global__ void k_ranges(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ misc,
                         int2* __restrict__ ranges) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t P = misc[1] ? 0 : (int64_t)misc[0];
  if (i >= P) return;
  const uint32_t t = keys[i];
  if (i == 0 || keys[i - 1] != t) ranges[t].x = (int)i;
  if (i == P - 1 || keys[i + 1] != t) ranges[t].y = (int)(i + 1);
}

// ------------------------------------------------------------------------------
// raster: per-Gaussian packing
// ------------------------------------------------------------------------------
template <int C>
__global__ void k_pack(const float2* __restrict__ means, const float* __restrict__ inv_cov,
                       const float* __restrict__ colours, const float* __restrict__ phase,
                       const float* __restrict__ opac, const float* __restrict__ curv,
                       const float* __restrict__ area, const float* __restrict__ kvec,
                       float dxy2, int n, float* __restrict__ packed) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const float2 m = means[g];
  const float beta = 0.5f * curv[g] * area[g] * dxy2;  // base = beta * mahal
  float r[REC];
  r[0] = m.x;
  r[1] = m.y;
  r[2] = QSCALE * inv_cov[3 * g + 0];
  r[3] = QSCALE * inv_cov[3 * g + 1];
  r[4] = QSCALE * inv_cov[3 * g + 2];
  r[5] = opac[g];
  r[6] = beta;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    r[7 + 3 * c] = phase[C * g + c] * INV_2PI;                   // u_c (revolutions)
    r[8 + 3 * c] = kvec[c] * beta * INV_2PI * (1.0f / QSCALE);    // w_c: revolutions per unit q
    r[9 + 3 * c] = colours[C * g + c];
  }
#pragma unroll
  for (int i = 7 + 3 * C; i < REC; ++i) r[i] = 0.f;
  float4* dst = reinterpret_cast<float4*>(packed + (size_t)g * REC);
#pragma unroll
  for (int i = 0; i < REC / 4; ++i) dst[i] = make_float4(r[4 * i], r[4 * i + 1], r[4 * i + 2], r[4 * i + 3]);
}

// Maximum over a tile's pixel rectangle [x0, x1] x [y0, y1] of the concave exponent
// q = a00 dx^2 + 2 a01 dx dy + a11 dy^2 (dx = mx - x, dy = my - y; the a's are the
// packed, log2-scaled, negative-definite coefficients): the centre if it lies inside,
// otherwise the largest of the four edge maxima.
__device__ __forceinline__ float tile_qmax(float mx, float my, float a00, float a01, float a11,
                                           float x0, float x1, float y0, float y1) {
  const float dxlo = mx - x1, dxhi = mx - x0, dylo = my - y1, dyhi = my - y0;
  if (dxlo <= 0.f && dxhi >= 0.f && dylo <= 0.f && dyhi >= 0.f) return 0.f;
  auto qf = [&](float dx, float dy) { return fmaf(dx, fmaf(a00, dx, 2.f * a01 * dy), a11 * dy * dy); };
  const float ry = -a01 / a11, rx = -a01 / a00;
  float best = qf(dxlo, fminf(fmaxf(ry * dxlo, dylo), dyhi));
  best = fmaxf(best, qf(dxhi, fminf(fmaxf(ry * dxhi, dylo), dyhi)));
  best = fmaxf(best, qf(fminf(fmaxf(rx * dylo, dxlo), dxhi), dylo));
  best = fmaxf(best, qf(fminf(fmaxf(rx * dyhi, dxlo), dxhi), dyhi));
  return best;
}

// Live count: k_count, but a tile of a Gaussian's rectangle is kept only if the Gaussian's
// exponent q (the packed, log2-scaled form the prologue wrote) can reach QSKIP on the
// tile's pixels.  Both raster kernels skip a key whose closed-form maximum over the tile
// (tile_qmax) is below QSKIP -- its terms are below 2^-24 of a unit term; the adjoint adds
// a zero partial record -- so leaving such keys out of the tile lists changes no result.
// Per row of tiles the test is one interval: the pixels x at which the maximum of q over
// the row's y range reaches thr = QSKIP - LIVE_MARGIN form an interval [xa, xb] (q is
// concave), bounded by the ellipse {q = thr}'s leftmost and rightmost points when they
// lie in the row and by its chords at the row's edges otherwise; a tile is kept if it
// meets [xa, xb] widened by 0.25 px + 1e-3 of the extent.  The lowered threshold and the
// widening keep every tile whose tile_qmax reaches QSKIP (checked on the CPU in float32
// over every key of the 1920x1080 / 150k and 3840x2160 / 600k scenes) at the cost of a
// few more keys (64.0 % against 63.4 % of the keys kept at 1080p).  live[g] marks the
// kept tiles, bit b for tile b of the rectangle in row-major order; rectangles of more
// than 64 tiles, and Gaussians whose packed record is not a finite negative-definite
// form, keep every tile.  counts and rect are those of k_count with the kept tiles.
// grid ceil(n / 256)
constexpr float LIVE_MARGIN = 0.0625f;

// The live count of one Gaussian from its mean, radius and packed exponent coefficients:
// the rectangle (lx, ly, width w), the kept tiles (bit mask) and their number.
__device__ __forceinline__ void live_tiles(float mx, float my, float a00, float a01, float a11,
                                           int r, int gx, int gy, int W, int H, int& lx,
                                           int& ly, int& w, uint32_t& c, uint64_t& m) {
  int hx = 0, hy = 0;
  lx = ly = 0;
  if (r > 0) get_rect(mx, my, r, gx, gy, lx, ly, hx, hy);
  w = hx - lx;
  const int area = (hx > lx && hy > ly) ? (hy - ly) * w : 0;
  c = 0;
  m = 0ull;
  if (area > 64) {
    c = (uint32_t)area;
    m = ~0ull;
    return;
  }
  if (area == 0) return;
  const float beta = a00 * a11 - a01 * a01;
  if (!(beta > 0.f) || !(a00 < 0.f) || !(a11 < 0.f) || !isfinite(mx) || !isfinite(my) ||
      !isfinite(beta)) {
    m = area == 64 ? ~0ull : (1ull << area) - 1ull;
  } else {
    // hardware reciprocals and square roots (1 ulp): far inside the widening below
    const float thr = QSKIP - LIVE_MARGIN;
    const float i00 = __builtin_amdgcn_rcpf(a00), ib = __builtin_amdgcn_rcpf(beta);
    const float rx = -a01 * i00, ry = -a01 * __builtin_amdgcn_rcpf(a11), ia = -i00;
    const float dxr = __builtin_amdgcn_sqrtf(fmaxf(thr * a11 * ib, 0.f));  // rightmost point
    const float dyr = ry * dxr;
    const float dymax = __builtin_amdgcn_sqrtf(fmaxf(a00 * thr * ib, 0.f));
    const float c0 = a00 * thr;
    for (int ty = ly; ty < hy; ++ty) {
      const float dylo = my - (float)(min(ty * TILE + TILE, H) - 1), dyhi = my - (float)(ty * TILE);
      if (dyhi < -dymax || dylo > dymax) continue;  // the row misses the ellipse
      // dx = mx - x of the interval's ends: the extreme points if the row holds them,
      // else the chord at the row edge nearest to them
      const float er = fminf(fmaxf(dyr, dylo), dyhi), el = fminf(fmaxf(-dyr, dylo), dyhi);
      const float dr = (dyr >= dylo && dyr <= dyhi)
          ? dxr : fmaf(rx, er, __builtin_amdgcn_sqrtf(fmaxf(c0 - beta * er * er, 0.f)) * ia);
      const float dl = (-dyr >= dylo && -dyr <= dyhi)
          ? -dxr : fmaf(rx, el, -__builtin_amdgcn_sqrtf(fmaxf(c0 - beta * el * el, 0.f)) * ia);
      const float xa = mx - (dr + (0.25f + 1e-3f * fabsf(dr)));
      const float xb = mx - (dl - (0.25f + 1e-3f * fabsf(dl)));
      const int b0 = (ty - ly) * w - lx;
      for (int tx = lx; tx < hx; ++tx) {
        const float x0 = (float)(tx * TILE), x1 = (float)(min(tx * TILE + TILE, W) - 1);
        if (x1 >= xa && x0 <= xb) m |= 1ull << (b0 + tx);
      }
    }
  }
  c = (uint32_t)__popcll(m);
}

__global__ __launch_bounds__(256) void k_count_live(
    const float2* __restrict__ means, const int32_t* __restrict__ radii,
    const float4* __restrict__ packed, int n, int gx, int gy, int W, int H,
    uint32_t* __restrict__ counts, int4* __restrict__ rect, uint64_t* __restrict__ live,
    uint32_t* __restrict__ misc) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0) misc[3] = 1u;
  if (g >= n) return;
  // every load up front: one memory latency
  const int r = radii[g];
  const float2 mm = means[g];
  const float4 q0 = packed[(size_t)g * (REC / 4)];
  const float a11 = packed[(size_t)g * (REC / 4) + 1].x;
  int lx, ly, w;
  uint32_t c;
  uint64_t m;
  live_tiles(mm.x, mm.y, q0.z, q0.w, a11, r, gx, gy, W, H, lx, ly, w, c, m);
  counts[g] = c;
  rect[g] = make_int4(0, lx, ly, c ? w : 0);
  live[g] = m;
}

// ------------------------------------------------------------------------------
// raster forward
// ------------------------------------------------------------------------------
// LDS record per Gaussian (16 floats): mx my a00 a01 | a11 u0 w0 amp0 | u1 w1 amp1 u2 |
// w2 amp2 direct -, with direct = 1 when the Gaussian's phases go to the sine units
// unreduced (C = 1: the first two float4s, and direct in the fourth)
// One 16x16 tile of one hologram; sg is the workgroup's 16 KiB staging buffer.  Every
// write to sg follows a workgroup barrier inside this function, so consecutive calls by
// one workgroup need no barrier between them.
template <int C>
__device__ __forceinline__ void render_fwd_tile(int tile, const int2* __restrict__ ranges,
                                                const int32_t* __restrict__ point_list,
                                                const float4* __restrict__ packed,
                                                const uint32_t* __restrict__ misc, int W, int H,
                                                int gx, float2* __restrict__ out, float4* sg,
                                                int* s_cnt) {
  const int tx = tile % gx, ty = tile / gx;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  // each wave covers an 8x8 quarter of the tile: more compact than four rows of 16, so
  // fewer (wave, Gaussian) pairs reach 2^-24 (44.5 % against 48.3 % of the tile lists'
  // pairs at 1080p, counted on the CPU from the synthetic scene)
  const int px = tx * TILE + ((t >> 6) & 1) * 8 + (t & 7);
  const int py = ty * TILE + (t >> 7) * 8 + ((t >> 3) & 7);
  int2 rg = ranges[tile];
  if (misc[1]) rg = make_int2(0, 0);
  const float fpx = (float)px, fpy = (float)py;
  const float tx0 = (float)(tx * TILE), ty0 = (float)(ty * TILE);
  const float tx1 = (float)(min(tx * TILE + TILE, W) - 1), ty1 = (float)(min(ty * TILE + TILE, H) - 1);
  float ar[C], ai[C];
#pragma unroll
  for (int c = 0; c < C; ++c) ar[c] = ai[c] = 0.f;
  for (int b0 = rg.x; b0 < rg.y; b0 += 256) {
    const int cnt = min(256, rg.y - b0);
    __syncthreads();
    // stage the batch's Gaussians, keeping (in key order) only those that reach
    // G = 2^-24 somewhere on the tile's image pixels: the others would be skipped by
    // every wave below
    float4 q0, q1, q2, q3;
    bool keep = false, direct = true;
    if (t < cnt) {
      const int g = point_list[b0 + t];
      const float4* src = packed + (size_t)g * (REC / 4);
      q0 = src[0]; q1 = src[1]; q2 = src[2]; q3 = src[3];
      keep = tile_qmax(q0.x, q0.y, q0.z, q0.w, q1.x, tx0, tx1, ty0, ty1) >= QSKIP;
      direct = rev_direct(q1.w, q2.x);
      if (C == 3) direct = direct && rev_direct(q2.z, q2.w) && rev_direct(q3.y, q3.z);
    }
    const uint64_t bal = __ballot(keep);
    if (lane == 0) s_cnt[wave] = __popcll(bal);
    // the batch's phases go to the sine units unreduced when every kept Gaussian allows it;
    // otherwise each Gaussian follows its own choice
    const bool reduce = __syncthreads_or(keep && !direct) != 0;
    int slot = __popcll(bal & ((1ull << lane) - 1ull));
    for (int w = 0; w < wave; ++w) slot += s_cnt[w];
    const int kept = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    if (keep) {
      // q0 = mx my a00 a01 ; q1 = a11 o beta u0 ; q2 = w0 col0 u1 w1 ; q3 = col1 u2 w2 col2
      const float o = q1.y;
      if (C == 3) {
        sg[4 * slot + 0] = q0;
        sg[4 * slot + 1] = make_float4(q1.x, q1.w, q2.x, q2.y * o);
        sg[4 * slot + 2] = make_float4(q2.z, q2.w, q3.x * o, q3.y);
        sg[4 * slot + 3] = make_float4(q3.z, q3.w * o, direct ? 1.f : 0.f, 0.f);
      } else {
        sg[4 * slot + 0] = q0;
        sg[4 * slot + 1] = make_float4(q1.x, q1.w, q2.x, q2.y * o);
        sg[4 * slot + 3] = make_float4(0.f, 0.f, direct ? 1.f : 0.f, 0.f);
      }
    }
    __syncthreads();
    auto walk = [&](auto fr) {
    constexpr bool FR = decltype(fr)::value;
    for (int j = 0; j < kept; ++j) {
      const float4 p0 = sg[4 * j + 0];
      const float4 p1 = sg[4 * j + 1];
      const float dx = p0.x - fpx, dy = p0.y - fpy;
      const float ex = fmaf(p0.w, dy, p0.z * dx);
      const float ey = fmaf(p1.x, dy, p0.w * dx);
      const float q = fmaf(dy, ey, dx * ex);
      // a wave whose 64 pixels all have G = 2^q < 2^-24 skips the Gaussian: each such
      // pair would add less than 2^-24 a o, below FP32 resolution of a unit-amplitude
      // term (the reference evaluates it, down to its clamp at e^-50); in other waves
      // pixels beyond the clamp get 2^q < e^-50 instead of the clamped value
      if (__ballot(q >= QSKIP) == 0ull) continue;
      const float G = fast_exp2(q);
      // the Gaussian's own choice, read only in a batch that mixes both kinds
      const bool dj = FR ? sg[4 * j + 3].z != 0.f : true;
      {
        const float rev = rev_of<FR>(fmaf(p1.z, q, p1.y), dj);
        const float a = p1.w * G;
        ar[0] = fmaf(a, cos_rev(rev), ar[0]);
        ai[0] = fmaf(a, sin_rev(rev), ai[0]);
      }
      if (C == 3) {
        const float4 p2 = sg[4 * j + 2];
        const float4 p3 = sg[4 * j + 3];
        const float rev1 = rev_of<FR>(fmaf(p2.y, q, p2.x), dj);
        const float a1 = p2.z * G;
        ar[1] = fmaf(a1, cos_rev(rev1), ar[1]);
        ai[1] = fmaf(a1, sin_rev(rev1), ai[1]);
        const float rev2 = rev_of<FR>(fmaf(p3.x, q, p2.w), dj);
        const float a2 = p3.y * G;
        ar[C - 1] = fmaf(a2, cos_rev(rev2), ar[C - 1]);
        ai[C - 1] = fmaf(a2, sin_rev(rev2), ai[C - 1]);
      }
    }
    };
    if (reduce) walk(std::true_type{});
    else walk(std::false_type{});
  }
  if (px < W && py < H) {
#pragma unroll
    for (int c = 0; c < C; ++c) out[((size_t)c * H + py) * W + px] = make_float2(ar[c], ai[c]);
  }
}

template <int C>
__global__ __launch_bounds__(256) void k_render_fwd(const int2* __restrict__ ranges,
                                                    const int32_t* __restrict__ point_list,
                                                    const float4* __restrict__ packed,
                                                    const uint32_t* __restrict__ misc, int W,
                                                    int H, int gx, float2* __restrict__ out) {
  __shared__ float4 sg[256 * 4];
  __shared__ int s_cnt[4];
  render_fwd_tile<C>(blockIdx.x, ranges, point_list, packed, misc, W, H, gx, out, sg, s_cnt);
}

// ------------------------------------------------------------------------------
// raster adjoint: one Gaussian per lane, each lane walking its own rows of the tile
// ------------------------------------------------------------------------------
// partial record (16 floats): dcol[C], dph'[C], dop, K', di00, di01', di11, dmx', dmy'
constexpr int GROW = TILE + 1;   // LDS row pitch of the pixel gradients (float2): rows that
                                 // lanes read at the same time fall in distinct bank pairs

// LDS of one adjoint workgroup
template <int C>
struct BwdLds {
  float2 grad[C][TILE * GROW];  // the tile's pixel gradients dL/dU, row pitch GROW
  int keys[512];                // queued keys, in key order
  int span[512];                // their rows: first significant row | row count << 8
  int rec[512];                 // their partial-record slots
  int rkeys[256];               // one evaluation round, by row count (descending)
  int rspan[256];
  int rrec[256];
  int cnt[4];
  int hist[TILE + 1];
};

// Lanes that evaluate one key together in the raster adjoint (2, 4 or 8).
#ifndef CVQPG_BWD_LANES
#define CVQPG_BWD_LANES 4
#endif
constexpr int BWD_LANES = CVQPG_BWD_LANES;
static_assert(BWD_LANES == 2 || BWD_LANES == 4 || BWD_LANES == 8, "2, 4 or 8 lanes per key");

// Sum over the BWD_LANES consecutive lanes of a key's group, the same value in each lane,
// in a fixed order ((v0 + v1) + (v2 + v3) for four lanes; for eight, the two quads' sums
// added), through DPP permutations.  Every lane of the group must be active.
__device__ __forceinline__ float group_sum(float v) {
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));  // quad_perm 1,0,3,2
  if (BWD_LANES >= 4)
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));  // quad_perm 2,3,0,1
  if (BWD_LANES == 8)
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, false));  // row_half_mirror
  return v;
}

// Rows of the tile on which a Gaussian reaches 2^QSKIP: for each row, the maximum of the
// exponent over the row's pixel span [x0, x1] (the concave quadratic's peak along the row,
// clamped to the span).  The row maximum is concave in y, so the rows form one range;
// returned as first | count << 8 (count 0: none).
__device__ __forceinline__ int row_span(float mx, float my, float a00, float a01, float a11,
                                        float x0, float x1, float y0, int ny) {
  const float dxlo = mx - x1, dxhi = mx - x0, rx = -a01 / a00;
  int first = TILE, last = -1;
#pragma unroll
  for (int yy = 0; yy < TILE; ++yy) {
    const float dy = my - (y0 + (float)yy);
    const float dx = fminf(fmaxf(rx * dy, dxlo), dxhi);
    const float q = fmaf(dx, fmaf(a00, dx, 2.f * a01 * dy), a11 * dy * dy);
    if (yy < ny && q >= QSKIP) {
      first = min(first, yy);
      last = yy;
    }
  }
  return last < first ? 0 : first | ((last - first + 1) << 8);
}

// Column blocks (NCOL = TILE / BWD_LANES columns each) of the tile in which a Gaussian
// reaches 2^QSKIP on the rows [yl, yh]: for each of the nx columns, the maximum of the
// exponent over those rows (the concave quadratic's peak along the column, clamped to the
// rows), the mirror of row_span.  Returned as first block | block count << 4 (the span's
// bits 16 and up); 0 if no column reaches QSKIP.
template <int NCOL>
__device__ __forceinline__ int col_blocks(float mx, float my, float a00, float a01, float a11,
                                          float x0, float yl, float yh, int nx) {
  const float dylo = my - yh, dyhi = my - yl, ry = -a01 / a11;
  int first = TILE, last = -1;
#pragma unroll
  for (int xx = 0; xx < TILE; ++xx) {
    const float dx = mx - (x0 + (float)xx);
    const float dy = fminf(fmaxf(ry * dx, dylo), dyhi);
    const float q = fmaf(dy, fmaf(a11, dy, 2.f * a01 * dx), a00 * dx * dx);
    if (xx < nx && q >= QSKIP) {
      first = min(first, xx);
      last = xx;
    }
  }
  if (last < first) return 0;
  const int b0 = first / NCOL;
  return b0 | ((last / NCOL - b0 + 1) << 4);
}

// Lane layout of a queued key in the raster adjoint, from its span (first row | row count
// << 8 | first column block << 16 | block count << 20).  By default lane cg of the key's
// group covers column block cg on every one of the key's rows.  When the key's significant
// columns lie in nb blocks and nb divides BWD_LANES (nb < BWD_LANES), the group's lanes
// cover those nb blocks only, BWD_LANES / nb lanes per block taking every
// (BWD_LANES / nb)-th row: block b0 + cg % nb, rows first + cg / nb, + stride, ...
// A key with no significant column (nb = 0) walks no rows.
struct KeyLanes {
  int blk, row0, stride, trips;   // this lane's block, first row, row stride, row count
};
__device__ __forceinline__ KeyLanes key_lanes(int span, int cg) {
  const int r0 = span & 0xff, nr = (span >> 8) & 0xff;
  const int b0 = (span >> 16) & 0xf, nb = (span >> 20) & 0xf;
  KeyLanes k{cg, r0, 1, nb == 0 ? 0 : nr};
  if (nb > 0 && nb < BWD_LANES && BWD_LANES % nb == 0) {
    k.stride = BWD_LANES / nb;
    k.blk = b0 + cg % nb;
    const int ro = cg / nb;
    k.row0 = r0 + ro;
    k.trips = nr > ro ? (nr - ro + k.stride - 1) / k.stride : 0;
  }
  return k;
}

// Row trips of a key's group: those of its lane 0, the most of any of its lanes.
__device__ __forceinline__ int key_trips(int span) { return key_lanes(span, 0).trips; }

// One tile of one hologram.  A key whose Gaussian stays below G = 2^-24 on every pixel of
// the tile adds less than 2^-24 of a unit term to every adjoint sum (the reference
// evaluates it, down to its clamp at e^-50): such keys get a zero partial record (so the
// per-Gaussian sums stay complete) and are not evaluated.  The others are queued with the
// range of rows on which they reach 2^-24 and evaluated in rounds of up to 256 keys, a
// group of BWD_LANES lanes per key, each group summing over the key's rows (rows outside
// the range add less than 2^-24 of a unit term, like the dropped keys).  Every round is
// first sorted by row count, so that the keys of a wave walk similar numbers of rows (a
// wave holds 64 / BWD_LANES keys, so the fewer lanes per key, the more the row counts
// within a wave differ); at 1080p the queued keys' row ranges hold about three quarters
// of their rows.  A key's record depends only on the key, so the order in which groups
// take keys does not change any result.  The first write to the
// pixel gradients is not preceded by a barrier here: a workgroup that calls this for
// several tiles places one between the calls.
template <int C>
__device__ __forceinline__ void render_bwd_tile(int tile, const int2* __restrict__ ranges,
                                                const int32_t* __restrict__ point_list,
                                                const float4* __restrict__ packed,
                                                const int4* __restrict__ rect,
                                                const uint64_t* __restrict__ live,
                                                const float2* __restrict__ grad,
                                                const float* __restrict__ kvec,
                                                const uint32_t* __restrict__ misc, int W, int H,
                                                int gx, float4* __restrict__ partials,
                                                BwdLds<C>& L) {
  const int tx = tile % gx, ty = tile / gx;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  int2 rg = ranges[tile];
  if (misc[1]) rg = make_int2(0, 0);
  if (rg.y <= rg.x) return;  // uniform
  {
    const int px = tx * TILE + (t & 15), py = ty * TILE + (t >> 4);
    const bool inside = px < W && py < H;
#pragma unroll
    for (int c = 0; c < C; ++c)
      L.grad[c][(t >> 4) * GROW + (t & 15)] =
          inside ? grad[((size_t)c * H + py) * W + px] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const int x0 = tx * TILE, y0 = ty * TILE;
  const int nx = min(TILE, W - x0), ny = min(TILE, H - y0);
  float kc[C];
#pragma unroll
  for (int c = 0; c < C; ++c) kc[c] = kvec[c];

  // BWD_LANES lanes per key: lane cg of the key's group covers the NCOL = 16 / BWD_LANES
  // columns NCOL cg .. NCOL cg + NCOL - 1 of the key's rows, and the group's sums are
  // combined by group_sum (a fixed order).  Columns past the image's right edge hold zero
  // pixel gradients, which add exact zeros to every sum.  Per column the lane keeps the
  // moments T0 = sum h and T1 = sum h dy over the rows, and T2 = sum h dy^2 over its
  // columns; the mean and inverse-covariance gradients are formed from them once per key.
  constexpr int NCOL = TILE / BWD_LANES;
  auto process = [&](int g, int span, int e, bool valid, int cg) {
    const float4* src = packed + (size_t)g * (REC / 4);
    const float4 q0 = src[0], q1 = src[1], q2 = src[2], q3 = src[3];
    const float mx = q0.x, my = q0.y, a00 = q0.z, a01 = q0.w, a11 = q1.x;
    const float o = q1.y, beta = q1.z;
    float u[C], w[C], col[C], kA[C], colh[C];
    u[0] = q1.w; w[0] = q2.x; col[0] = q2.y;
    if (C == 3) {
      u[1] = q2.z; w[1] = q2.w; col[1] = q3.x;
      u[C - 1] = q3.y; w[C - 1] = q3.z; col[C - 1] = q3.w;
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
      kA[c] = kc[c] * col[c];
      colh[c] = -0.5f * col[c];
    }
    float dcol[C], dph[C];
#pragma unroll
    for (int c = 0; c < C; ++c) dcol[c] = dph[c] = 0.f;
    float K = 0.f, T2 = 0.f, T0[NCOL], T1[NCOL];
#pragma unroll
    for (int j = 0; j < NCOL; ++j) T0[j] = T1[j] = 0.f;
    const KeyLanes kl = key_lanes(span, cg);
    const float dxc = mx - (float)(x0 + NCOL * kl.blk);
    // the round is sorted by row trips, so the wave's first lane walks the most rows
    const int nmax = __builtin_amdgcn_readfirstlane(kl.trips);
    bool direct = true;
#pragma unroll
    for (int c = 0; c < C; ++c) direct = direct && rev_direct(u[c], w[c]);
    // the wave's phases go to the sine units unreduced when all its keys allow it;
    // otherwise each key follows its own choice
    const bool reduce = __ballot(valid && !direct) != 0ull;
    auto walk = [&](auto fr) {
    constexpr bool FR = decltype(fr)::value;
    for (int k = 0; k < nmax; ++k) {
      if (k >= kl.trips) continue;
      const int yy = kl.row0 + k * kl.stride;
      const float dy = my - (float)(y0 + yy);
      const float bq = 2.0f * a01 * dy, cq = a11 * dy * dy;
      const float2* gr0 = &L.grad[0][yy * GROW + NCOL * kl.blk];
#pragma unroll
      for (int j = 0; j < NCOL; ++j) {
        const float dx = dxc - (float)j;
        const float q = fmaf(dx, fmaf(a00, dx, bq), cq);
        const float G = fast_exp2(q);
        float Sh = 0.f, Tb = 0.f;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const float2 gg = gr0[c * TILE * GROW + j];
          const float rev = rev_of<FR>(fmaf(w[c], q, u[c]), direct);
          const float sn = sin_rev(rev), cs = cos_rev(rev);
          const float Pc = fmaf(gg.x, cs, gg.y * sn);
          const float Qc = fmaf(gg.y, cs, -gg.x * sn);
          dcol[c] = fmaf(G, Pc, dcol[c]);
          dph[c] = fmaf(G, Qc, dph[c]);
          Sh = fmaf(colh[c], Pc, Sh);
          Tb = fmaf(kA[c], Qc, Tb);
        }
        const float h = G * fmaf(Tb, beta, Sh);
        K = fmaf(G * Tb, q, K);
        T0[j] += h;
        const float hy = h * dy;
        T1[j] += hy;
        T2 = fmaf(hy, dy, T2);
      }
    }
    };
    if (reduce) walk(std::true_type{});
    else walk(std::false_type{});
    // S1 = sum h dx, M00 = sum h dx^2, M01 = sum h dx dy, M1y = sum h dy over the lane
    float S1 = 0.f, M00 = 0.f, M01 = 0.f, M1y = 0.f;
#pragma unroll
    for (int j = 0; j < NCOL; ++j) {
      const float dx = dxc - (float)j;
      const float t0 = T0[j] * dx;
      S1 += t0;
      M00 = fmaf(t0, dx, M00);
      M01 = fmaf(T1[j], dx, M01);
      M1y += T1[j];
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
      dcol[c] = group_sum(dcol[c]);
      dph[c] = group_sum(dph[c]);
    }
    K = group_sum(K);
    S1 = group_sum(S1);
    M00 = group_sum(M00);
    M01 = group_sum(M01);
    M1y = group_sum(M1y);
    T2 = group_sum(T2);
    float dop = 0.f;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      dop = fmaf(col[c], dcol[c], dop);
      dcol[c] *= o;
      dph[c] *= o;
    }
    if (valid) {
      float rec[REC];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        rec[c] = dcol[c];
        rec[C + c] = dph[c];
      }
      rec[2 * C + 0] = dop;
      rec[2 * C + 1] = K * o;
      rec[2 * C + 2] = M00 * o;
      rec[2 * C + 3] = M01 * o;
      rec[2 * C + 4] = T2 * o;
      rec[2 * C + 5] = fmaf(a00, S1, a01 * M1y) * o;
      rec[2 * C + 6] = fmaf(a01, S1, a11 * M1y) * o;
#pragma unroll
      for (int i = 2 * C + 7; i < REC; ++i) rec[i] = 0.f;
      // the group's lanes write the record's float4s (lane cg: numbers cg, cg + BWD_LANES,
      // ...): one 64-byte run
#pragma unroll
      for (int i0 = 0; i0 < REC / 4; i0 += BWD_LANES) {
        float4 v = make_float4(rec[4 * i0], rec[4 * i0 + 1], rec[4 * i0 + 2], rec[4 * i0 + 3]);
#pragma unroll
        for (int i = i0 + 1; i < i0 + BWD_LANES && i < REC / 4; ++i)
          if (cg == i - i0) v = make_float4(rec[4 * i], rec[4 * i + 1], rec[4 * i + 2], rec[4 * i + 3]);
        if (i0 + cg < REC / 4) partials[(size_t)e * (REC / 4) + i0 + cg] = v;
      }
    }
  };

  const float fx0 = (float)x0, fy0 = (float)y0;
  const float fx1 = (float)(x0 + nx - 1), fy1 = (float)(y0 + ny - 1);
  // a key's partial record goes to its Gaussian's emission slot for this tile: the tile's
  // place in the rectangle, or after a live count (misc[3]) the number of kept tiles
  // before it (rectangles of more than 64 tiles keep every tile: live[g] is all ones)
  const bool use_live = live != nullptr && misc[3] != 0u;
  int filled = 0;  // queued keys in L.keys[0, filled)
  for (int base = rg.x; base < rg.y; base += 256) {
    const int i = base + t;
    bool keep = false;
    int g = 0, span = 0, e = 0;
    if (i < rg.y) {
      g = point_list[i];
      const int4 rc = rect[g];
      int b = (ty - rc.z) * rc.w + (tx - rc.y);
      if (use_live && b < 64) b = __popcll(live[g] & ((1ull << b) - 1ull));
      e = rc.x + b;
      const float4 q0 = packed[(size_t)g * (REC / 4)];
      const float a11 = packed[(size_t)g * (REC / 4) + 1].x;
      keep = tile_qmax(q0.x, q0.y, q0.z, q0.w, a11, fx0, fx1, fy0, fy1) >= QSKIP;
      if (keep) {
        span = row_span(q0.x, q0.y, q0.z, q0.w, a11, fx0, fx1, fy0, ny);
        const int r0 = span & 0xff, nr = span >> 8;
        if (nr > 0)
          span |= col_blocks<NCOL>(q0.x, q0.y, q0.z, q0.w, a11, fx0, fy0 + (float)r0,
                                   fy0 + (float)(r0 + nr - 1), nx) << 16;
      }
      if (!keep) {
        float4* dst = partials + (size_t)e * (REC / 4);
#pragma unroll
        for (int k = 0; k < REC / 4; ++k) dst[k] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
    const uint64_t bal = __ballot(keep);
    if (lane == 0) L.cnt[wave] = __popcll(bal);
    __syncthreads();
    int pos = filled + __popcll(bal & ((1ull << lane) - 1ull));
    for (int w = 0; w < wave; ++w) pos += L.cnt[w];
    if (keep) {
      L.keys[pos] = g;
      L.span[pos] = span;
      L.rec[pos] = e;
    }
    filled += L.cnt[0] + L.cnt[1] + L.cnt[2] + L.cnt[3];
    __syncthreads();  // the queue is complete; cnt may be rewritten from here on
    const bool last = base + 256 >= rg.y;
    while (filled >= 256 || (last && filled > 0)) {
      const int n = min(filled, 256);
      // counting sort of the round's n keys by row trips, most first; the order within
      // one count is that of the LDS atomics and does not matter
      if (t <= TILE) L.hist[t] = 0;
      __syncthreads();
      int rk = 0, sp = 0, key = 0, er = 0;
      if (t < n) {
        key = L.keys[t];
        sp = L.span[t];
        er = L.rec[t];
        rk = atomicAdd(&L.hist[TILE - key_trips(sp)], 1);
      }
      __syncthreads();
      if (t == 0) {
        int acc = 0;
        for (int b = 0; b <= TILE; ++b) {
          const int c = L.hist[b];
          L.hist[b] = acc;
          acc += c;
        }
      }
      __syncthreads();
      if (t < n) {
        const int at = L.hist[TILE - key_trips(sp)] + rk;
        L.rkeys[at] = key;
        L.rspan[at] = sp;
        L.rrec[at] = er;
      }
      if (t < filled - 256) {
        L.keys[t] = L.keys[256 + t];
        L.span[t] = L.span[256 + t];
        L.rec[t] = L.rec[256 + t];
      }
      filled -= n;
      __syncthreads();
      // groups of KPW consecutive keys of the sorted round, BWD_LANES lanes per key; wave w
      // takes the groups w, w + 4, w + 8, ... so that every wave gets long and short keys
      constexpr int KPW = 64 / BWD_LANES;
      for (int gi = wave; gi * KPW < n; gi += 4) {
        const int slot = gi * KPW + lane / BWD_LANES;
        process(L.rkeys[slot < n ? slot : 0], slot < n ? L.rspan[slot] : 0,
                L.rrec[slot < n ? slot : 0], slot < n, lane % BWD_LANES);
      }
      __syncthreads();  // every lane is done with rkeys / rspan
    }
  }
}

#ifndef CVQPG_BWD_WPE
#define CVQPG_BWD_WPE 5
#endif
template <int C>
__global__ __launch_bounds__(256, CVQPG_BWD_WPE) void k_render_bwd(const int2* __restrict__ ranges,
                                                    const int32_t* __restrict__ point_list,
                                                    const float4* __restrict__ packed,
                                                    const int4* __restrict__ rect,
                                                    const uint64_t* __restrict__ live,
                                                    const float2* __restrict__ grad,
                                                    const float* __restrict__ kvec,
                                                    const uint32_t* __restrict__ misc, int W,
                                                    int H, int gx, float4* __restrict__ partials) {
  __shared__ BwdLds<C> lds;
  render_bwd_tile<C>(blockIdx.x, ranges, point_list, packed, rect, live, grad, kvec, misc, W,
                     H, gx, partials, lds);
}

// ------------------------------------------------------------------------------
// many holograms in one launch: a persistent grid drains a queue of tiles
// ------------------------------------------------------------------------------
// The queue is the concatenation of every scene's tiles in table order.  A workgroup
// takes `chunk` consecutive tiles per atomic fetch until the queue is empty; the last
// workgroup to leave resets the fetch and exit counters to zero, so the next launch
// (stream-ordered after this one) starts from a clean queue with no separate clear.
constexpr int MAX_BATCH = 64;

struct BatchScene {
  const int2* ranges;
  const int32_t* point_list;
  const float4* packed;
  const int4* rect;
  const uint32_t* misc;
  float2* field;
  const float2* grad;
  const float* k;
  float4* partials;
  const uint64_t* live;  // kept tiles of each rectangle after a live count
  int W, H, gx, tile_begin;
};

struct BatchHeader {
  uint32_t fetch[2];  // [0] forward, [1] adjoint
  uint32_t exits[2];
  int n_scenes, total_tiles, pad[10];
};
static_assert(sizeof(BatchHeader) == 64, "batch header is one 64-byte line");

__device__ __forceinline__ int batch_next(BatchHeader* hdr, int which, int chunk, int* slot) {
  __syncthreads();  // every lane has read the previous slot value and left the previous tile
  if (threadIdx.x == 0) *slot = (int)atomicAdd(&hdr->fetch[which], 1u) * chunk;
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(*slot);
}

__device__ __forceinline__ void batch_exit(BatchHeader* hdr, int which) {
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&hdr->exits[which], 1u) == gridDim.x - 1) {
      atomicExch(&hdr->fetch[which], 0u);
      atomicExch(&hdr->exits[which], 0u);
    }
  }
}

__device__ __forceinline__ int batch_scene_of(const int* begin, int ns, int item, int s) {
  while (s + 1 < ns && item >= begin[s + 1]) ++s;
  return __builtin_amdgcn_readfirstlane(s);
}

template <int C>
__global__ __launch_bounds__(256) void k_render_fwd_batch(BatchHeader* __restrict__ hdr,
                                                          const BatchScene* __restrict__ sc,
                                                          int chunk) {
  __shared__ float4 sg[256 * 4];
  __shared__ int begin[MAX_BATCH];
  __shared__ int slot;
  __shared__ int s_cnt[4];
  const int ns = hdr->n_scenes, total = hdr->total_tiles;
  for (int i = threadIdx.x; i < ns; i += 256) begin[i] = sc[i].tile_begin;
  for (;;) {
    const int first = batch_next(hdr, 0, chunk, &slot);
    if (first >= total) break;
    const int last = min(first + chunk, total);
    int s = 0;
    for (int it = first; it < last; ++it) {
      s = batch_scene_of(begin, ns, it, s);
      const BatchScene& S = sc[s];
      render_fwd_tile<C>(it - S.tile_begin, S.ranges, S.point_list, S.packed, S.misc, S.W, S.H,
                         S.gx, S.field, sg, s_cnt);
    }
  }
  batch_exit(hdr, 0);
}

template <int C>
__global__ __launch_bounds__(256, CVQPG_BWD_WPE) void k_render_bwd_batch(BatchHeader* __restrict__ hdr,
                                                          const BatchScene* __restrict__ sc,
                                                          int chunk) {
  __shared__ BwdLds<C> lds;
  __shared__ int begin[MAX_BATCH];
  __shared__ int slot;
  const int ns = hdr->n_scenes, total = hdr->total_tiles;
  for (int i = threadIdx.x; i < ns; i += 256) begin[i] = sc[i].tile_begin;
  for (;;) {
    const int first = batch_next(hdr, 1, chunk, &slot);
    if (first >= total) break;
    const int last = min(first + chunk, total);
    int s = 0;
    for (int it = first; it < last; ++it) {
      if (it > first) __syncthreads();  // the previous tile's waves are done with lds
      s = batch_scene_of(begin, ns, it, s);
      const BatchScene& S = sc[s];
      render_bwd_tile<C>(it - S.tile_begin, S.ranges, S.point_list, S.packed, S.rect, S.live,
                         S.grad, S.k, S.misc, S.W, S.H, S.gx, S.partials, lds);
    }
  }
  batch_exit(hdr, 1);
}

// per-Gaussian sum of the partial records over its contiguous emission slots
template <int C>
__global__ void k_bwd_reduce(const float4* __restrict__ partials, const int4* __restrict__ rect,
                             const uint32_t* __restrict__ counts, const uint32_t* __restrict__ misc,
                             const float* __restrict__ colours, const float* __restrict__ curv,
                             const float* __restrict__ area, float dxy2, int n,
                             float* __restrict__ d_means, float* __restrict__ d_inv,
                             float* __restrict__ d_col, float* __restrict__ d_ph,
                             float* __restrict__ d_op, float* __restrict__ d_cv,
                             float* __restrict__ d_ar) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  float acc[REC];
#pragma unroll
  for (int i = 0; i < REC; ++i) acc[i] = 0.f;
  if (!misc[1]) {
    const int e0 = rect[g].x;
    const int cnt = (int)counts[g];
    for (int e = e0; e < e0 + cnt; ++e) {
      const float4* src = partials + (size_t)e * (REC / 4);
#pragma unroll
      for (int i = 0; i < REC / 4; ++i) {
        const float4 v = src[i];
        acc[4 * i] += v.x;
        acc[4 * i + 1] += v.y;
        acc[4 * i + 2] += v.z;
        acc[4 * i + 3] += v.w;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    d_col[C * g + c] = acc[c];
    d_ph[C * g + c] = acc[C + c] * colours[C * g + c];
  }
  const float Kq = acc[2 * C + 1] * (1.0f / QSCALE);  // sum alpha * Tb * mahal
  d_op[g] = acc[2 * C + 0];
  d_cv[g] = Kq * 0.5f * area[g] * dxy2;
  d_ar[g] = Kq * 0.5f * curv[g] * dxy2;
  d_inv[3 * g + 0] = acc[2 * C + 2];
  d_inv[3 * g + 1] = 2.0f * acc[2 * C + 3];
  d_inv[3 * g + 2] = acc[2 * C + 4];
  d_means[2 * g + 0] = acc[2 * C + 5] * (2.0f / QSCALE);
  d_means[2 * g + 1] = acc[2 * C + 6] * (2.0f / QSCALE);
}

// ------------------------------------------------------------------------------
// fused prologue (activations -> covariance -> inverse -> radius -> packed record) and
// the adjoint reduction chained back to the seven raw parameters
//   model_2d_gaussian_rgb.py:83-153 (activations, covariance, inverse) and :428-436
//   (inv_cov stack, area, 5-sigma radius)
// ------------------------------------------------------------------------------
constexpr int NPARAM = 7;  // means_2d, pre_act_scales, pre_act_rotation, colours,
                           // pre_act_phase, pre_act_opacities, curvature
struct RawPtrs {
  const float* p[NPARAM];
};

struct Prologue {
  int W, H;
  float dxy2;
  float curv_lambda;   // wavelength of the curvature bound (min over channels; first for gray)
  float curv_nyquist;  // scale-aware bound fraction (0.9)
  int curv_mode;       // 0 off, 1 constant (200 tanh), 2 scale-aware
};

template <int C>
struct Act {
  float tmx, tmy, mx, my, sx, sy, rot, cs, sn, o, tk, cmax, curv;
  float c00, c01, c11, det, i00, i01, i11, area;
  bool clamped;
  float col[C], ph[C];
};

template <int C>
__device__ __forceinline__ void activate(const RawPtrs& R, int g, const Prologue& pr, Act<C>& a) {
  const float2 m = reinterpret_cast<const float2*>(R.p[0])[g];
  a.tmx = tanhf(m.x);
  a.tmy = tanhf(m.y);
  a.mx = (a.tmx + 1.0f) * 0.5f * (float)pr.W;
  a.my = (a.tmy + 1.0f) * 0.5f * (float)pr.H;
  const float2 s = reinterpret_cast<const float2*>(R.p[1])[g];
  a.sx = expf(s.x) + 0.1f;
  a.sy = expf(s.y) + 0.1f;
  a.rot = R.p[2][g];
  constexpr float TWO_PI = 6.283185307179586f;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    a.col[c] = R.p[3][C * g + c];
    const float x = R.p[4][C * g + c];
    float r = fmodf(x, TWO_PI);  // Python-style remainder (sign of the divisor)
    if (r < 0.f) r += TWO_PI;
    if (r >= TWO_PI) r -= TWO_PI;
    a.ph[c] = r;
  }
  a.o = 1.0f / (1.0f + expf(-R.p[5][g]));
  const float sig = fmaxf(a.sx, a.sy);
  a.tk = tanhf(R.p[6][g]);
  if (pr.curv_mode == 2) {
    a.cmax = pr.curv_nyquist * pr.curv_lambda / (6.0f * sig * pr.dxy2 + 1e-20f);
    a.curv = a.cmax * a.tk;
  } else if (pr.curv_mode == 1) {
    a.cmax = 200.0f;
    a.curv = 200.0f * a.tk;
  } else {
    a.cmax = 0.f;
    a.curv = 0.f;
  }
  a.cs = cosf(a.rot);
  a.sn = sinf(a.rot);
  const float sx2 = a.sx * a.sx, sy2 = a.sy * a.sy;
  a.c00 = sx2 * a.cs * a.cs + sy2 * a.sn * a.sn + 0.1f;
  a.c01 = (sx2 - sy2) * a.cs * a.sn;
  a.c11 = sx2 * a.sn * a.sn + sy2 * a.cs * a.cs + 0.1f;
  const float det = a.c00 * a.c11 - a.c01 * a.c01;
  a.clamped = !(det >= 1e-10f);
  a.det = a.clamped ? 1e-10f : det;
  const float inv_det = 1.0f / a.det;
  a.i00 = a.c11 * inv_det;
  a.i01 = -a.c01 * inv_det;
  a.i11 = a.c00 * inv_det;
  a.area = fabsf(a.sx * a.sy);
}

// Bins (LIVE = true: also the live count of k_count_live, from the values in registers;
// counts / rect / live / misc as k_count_live writes them, bit for bit).
struct LiveOut {
  int gx, gy;
  uint32_t* counts;
  int4* rect;
  uint64_t* live;
  uint32_t* misc;
};

template <int C, bool LIVE>
__global__ void k_prologue(RawPtrs R, Prologue pr, const float* __restrict__ kvec, int n,
                           float* __restrict__ means_out, int32_t* __restrict__ radii_out,
                           float* __restrict__ packed, LiveOut lo) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (LIVE && g == 0) lo.misc[3] = 1u;
  if (g >= n) return;
  Act<C> a;
  activate<C>(R, g, pr, a);
  reinterpret_cast<float2*>(means_out)[g] = make_float2(a.mx, a.my);
  const float mid = 0.5f * (a.c00 + a.c11);
  const float disc = sqrtf(((a.c00 - a.c11) * 0.5f) * ((a.c00 - a.c11) * 0.5f) + a.c01 * a.c01);
  const int32_t rad = (int32_t)ceilf(5.0f * sqrtf(mid + disc));
  radii_out[g] = rad;
  const float beta = 0.5f * a.curv * a.area * pr.dxy2;
  float r[REC];
  r[0] = a.mx;
  r[1] = a.my;
  r[2] = QSCALE * a.i00;
  r[3] = QSCALE * a.i01;
  r[4] = QSCALE * a.i11;
  r[5] = a.o;
  r[6] = beta;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    r[7 + 3 * c] = a.ph[c] * INV_2PI;
    r[8 + 3 * c] = kvec[c] * beta * INV_2PI * (1.0f / QSCALE);
    r[9 + 3 * c] = a.col[c];
  }
#pragma unroll
  for (int i = 7 + 3 * C; i < REC; ++i) r[i] = 0.f;
  float4* dst = reinterpret_cast<float4*>(packed + (size_t)g * REC);
#pragma unroll
  for (int i = 0; i < REC / 4; ++i) dst[i] = make_float4(r[4 * i], r[4 * i + 1], r[4 * i + 2], r[4 * i + 3]);
  if (LIVE) {
    int lx, ly, w;
    uint32_t c;
    uint64_t m;
    live_tiles(r[0], r[1], r[2], r[3], r[4], rad, lo.gx, lo.gy, pr.W, pr.H, lx, ly, w, c, m);
    lo.counts[g] = c;
    lo.rect[g] = make_int4(0, lx, ly, c ? w : 0);
    lo.live[g] = m;
  }
}

__device__ __forceinline__ float finite_or_zero(float v) { return isfinite(v) ? v : 0.0f; }

// grads: flat [(7 + 2C) N] in parameter order (means [N,2], scales [N,2], rotation [N],
// colours [N,C], phase [N,C], opacities [N], curvature [N]); sumsq: one partial per block
template <int C>
__global__ __launch_bounds__(256) void k_bwd_reduce_raw(
    const float4* __restrict__ partials, const int4* __restrict__ rect,
    const uint32_t* __restrict__ counts, const uint32_t* __restrict__ misc, RawPtrs R,
    Prologue pr, const int32_t* __restrict__ itr, int curv_warmup, int n,
    float* __restrict__ grads, float* __restrict__ sumsq) {
  __shared__ float red[256];
  __shared__ float s_acc[256][REC + 1];
  // 1. sums of the partial records, in key order: four lanes per Gaussian, lane q adding
  //    floats 4q .. 4q + 3 of each record (a wave load covers 16 Gaussians' 64-byte
  //    records), four records in flight per lane; 64 Gaussians per pass, four passes
#pragma unroll 1
  for (int pass = 0; pass < 4; ++pass) {
    const int lg = pass * 64 + (threadIdx.x >> 2), q = threadIdx.x & 3;
    const int gg = blockIdx.x * blockDim.x + lg;
    float4 a4 = make_float4(0.f, 0.f, 0.f, 0.f);
    if (gg < n && !misc[1]) {
      const int e0 = rect[gg].x;
      const int e1 = e0 + (int)counts[gg];
      for (int e = e0; e < e1; e += 4) {
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
          v[u] = e + u < e1 ? partials[(size_t)(e + u) * (REC / 4) + q]
                            : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (e + u < e1) {
            a4.x += v[u].x;
            a4.y += v[u].y;
            a4.z += v[u].z;
            a4.w += v[u].w;
          }
        }
      }
    }
    s_acc[lg][4 * q] = a4.x;
    s_acc[lg][4 * q + 1] = a4.y;
    s_acc[lg][4 * q + 2] = a4.z;
    s_acc[lg][4 * q + 3] = a4.w;
  }
  __syncthreads();
  // 2. one thread per Gaussian from here on
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  float sq = 0.f;
  if (g < n) {
    float acc[REC];
#pragma unroll
    for (int i = 0; i < REC; ++i) acc[i] = s_acc[threadIdx.x][i];
    Act<C> a;
    activate<C>(R, g, pr, a);
    // gradients with respect to the rasterizer inputs
    const float Kq = acc[2 * C + 1] * (1.0f / QSCALE);
    const float d_op = acc[2 * C + 0];
    const float d_cv = Kq * 0.5f * a.area * pr.dxy2;
    const float d_ar = Kq * 0.5f * a.curv * pr.dxy2;
    const float d00 = acc[2 * C + 2], d01 = 2.0f * acc[2 * C + 3], d11 = acc[2 * C + 4];
    const float dmx = acc[2 * C + 5] * (2.0f / QSCALE), dmy = acc[2 * C + 6] * (2.0f / QSCALE);
    // inverse <- covariance
    float gc00, gc01, gc11;
    if (a.clamped) {
      const float id = 1.0f / a.det;
      gc00 = d11 * id;
      gc11 = d00 * id;
      gc01 = -d01 * id;
    } else {
      const float D = a.det, iD = 1.0f / D, iD2 = iD * iD;
      gc00 = -d00 * a.c11 * a.c11 * iD2 + d01 * a.c01 * a.c11 * iD2 + d11 * (iD - a.c00 * a.c11 * iD2);
      gc11 = d00 * (iD - a.c11 * a.c00 * iD2) + d01 * a.c01 * a.c00 * iD2 - d11 * a.c00 * a.c00 * iD2;
      gc01 = d00 * 2.0f * a.c11 * a.c01 * iD2 + d01 * (-iD - 2.0f * a.c01 * a.c01 * iD2) +
             d11 * 2.0f * a.c00 * a.c01 * iD2;
    }
    // covariance <- scales, rotation
    const float c2 = a.cs * a.cs, s2 = a.sn * a.sn, csn = a.cs * a.sn;
    float gsx = 2.0f * a.sx * (gc00 * c2 + gc01 * csn + gc11 * s2);
    float gsy = 2.0f * a.sy * (gc00 * s2 - gc01 * csn + gc11 * c2);
    const float grot = (a.sx * a.sx - a.sy * a.sy) *
                       (-2.0f * csn * gc00 + (c2 - s2) * gc01 + 2.0f * csn * gc11);
    // area = |sx sy|
    const float sgn = (a.sx * a.sy) >= 0.f ? 1.0f : -1.0f;
    gsx += d_ar * sgn * a.sy;
    gsy += d_ar * sgn * a.sx;
    // curvature = cmax(sigma) tanh(kappa)
    float gk = 0.f;
    if (pr.curv_mode == 2) {
      const float sig = fmaxf(a.sx, a.sy);
      gk = d_cv * a.cmax * (1.0f - a.tk * a.tk);
      const float dcm = -a.cmax * 6.0f * pr.dxy2 / (6.0f * sig * pr.dxy2 + 1e-20f);
      const float gsig = d_cv * a.tk * dcm;
      if (a.sx >= a.sy) gsx += gsig; else gsy += gsig;
    } else if (pr.curv_mode == 1) {
      gk = d_cv * 200.0f * (1.0f - a.tk * a.tk);
    }
    const bool curv_on = pr.curv_mode != 0 && (itr ? itr[0] : 0) >= curv_warmup;
    // raw parameters
    const int N = n;
    float* gm = grads;
    float* gs = grads + 2 * (size_t)N;
    float* gr = grads + 4 * (size_t)N;
    float* gcol = grads + 5 * (size_t)N;
    float* gph = grads + (5 + C) * (size_t)N;
    float* gop = grads + (5 + 2 * C) * (size_t)N;
    float* gcv = grads + (6 + 2 * C) * (size_t)N;
    float v;
    v = finite_or_zero(dmx * 0.5f * (float)pr.W * (1.0f - a.tmx * a.tmx)); gm[2 * g] = v; sq += v * v;
    v = finite_or_zero(dmy * 0.5f * (float)pr.H * (1.0f - a.tmy * a.tmy)); gm[2 * g + 1] = v; sq += v * v;
    v = finite_or_zero(gsx * (a.sx - 0.1f)); gs[2 * g] = v; sq += v * v;
    v = finite_or_zero(gsy * (a.sy - 0.1f)); gs[2 * g + 1] = v; sq += v * v;
    v = finite_or_zero(grot); gr[g] = v; sq += v * v;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      v = finite_or_zero(acc[c]); gcol[C * g + c] = v; sq += v * v;
      v = finite_or_zero(acc[C + c] * a.col[c]); gph[C * g + c] = v; sq += v * v;
    }
    v = finite_or_zero(d_op * a.o * (1.0f - a.o)); gop[g] = v; sq += v * v;
    v = curv_on ? finite_or_zero(gk) : 0.0f; gcv[g] = v; sq += v * v;
  }
  red[threadIdx.x] = sq;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) sumsq[blockIdx.x] = red[0];
}

// ------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------
static int tile_bits(int T) {
  int b = 0;
  while ((1 << b) < T) ++b;
  return b;
}

static int check_common(int n, int channels, int width, int height, void* ws, size_t ws_bytes,
                        int64_t capacity) {
  CVQPG_CHECK_ARG(n >= 0, "n must be >= 0 (got %d)", n);
  CVQPG_CHECK_ARG(channels == 1 || channels == 3,
                  "qpf rasterizer supports 1 or 3 channels, got %d", channels);
  CVQPG_CHECK_ARG(width > 0 && height > 0, "invalid image size %dx%d", width, height);
  CVQPG_CHECK_ARG(capacity >= 0 && capacity < (int64_t)1 << 31, "invalid capacity %lld",
                  (long long)capacity);
  CVQPG_CHECK_ARG(ws != nullptr, "workspace is NULL");
  const size_t need = carve(nullptr, n, capacity).bytes;
  CVQPG_CHECK_ARG(ws_bytes >= need, "workspace too small: %zu < %zu bytes", ws_bytes, need);
  return 0;
}

}  // namespace cvqpg

using namespace cvqpg;

extern "C" size_t cvqpg_raster_workspace_size(int n, int channels, int width, int height,
                                              int64_t capacity) {
  (void)channels;
  (void)width;
  (void)height;
  return carve(nullptr, n, capacity).bytes;
}

extern "C" int cvqpg_raster_overflow_flags(void* ws, size_t ws_bytes, int n, int64_t capacity,
                                           uint32_t** flags) {
  if (int rc = check_common(n, 1, 1, 1, ws, ws_bytes, capacity)) return rc;
  CVQPG_CHECK_ARG(flags != nullptr, "flags is NULL");
  *flags = carve(ws, n, capacity).misc + 1;  // host arithmetic only: no device access
  return 0;
}

extern "C" int cvqpg_raster_clear_overflow(void* ws, size_t ws_bytes, int n, int64_t capacity,
                                           void* stream) {
  if (int rc = check_common(n, 1, 1, 1, ws, ws_bytes, capacity)) return rc;
  hipLaunchKernelGGL(k_clear_overflow, dim3(1), dim3(64), 0, (hipStream_t)stream,
                     carve(ws, n, capacity).misc);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_bin_count(const float* means, const int32_t* radii, int n, int width,
                               int height, void* ws, size_t ws_bytes, int64_t capacity,
                               int32_t* num_pairs, void* stream) {
  if (int rc = check_common(n, 1, width, height, ws, ws_bytes, capacity)) return rc;
  hipStream_t s = (hipStream_t)stream;
  RasterWS w = carve(ws, n, capacity);
  const int gx = ceil_div(width, TILE), gy = ceil_div(height, TILE);
  if (n > 0) {
    CVQPG_CHECK_ARG(means && radii, "means/radii are NULL");
    hipLaunchKernelGGL(k_count, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                       reinterpret_cast<const float2*>(means), radii, n, gx, gy, w.counts, w.rect,
                       w.misc);
    CVQPG_LAUNCH_CHECK();
  }
  return scan_exclusive(w.counts, w.offs, w.sums_n, n, w.misc, capacity, num_pairs, nullptr, s);
}

extern "C" int cvqpg_bin_count_live(const float* means, const int32_t* radii, int n, int width,
                                    int height, void* ws, size_t ws_bytes, int64_t capacity,
                                    int32_t* num_pairs, void* stream) {
  if (int rc = check_common(n, 1, width, height, ws, ws_bytes, capacity)) return rc;
  hipStream_t s = (hipStream_t)stream;
  RasterWS w = carve(ws, n, capacity);
  const int gx = ceil_div(width, TILE), gy = ceil_div(height, TILE);
  if (n > 0) {
    CVQPG_CHECK_ARG(means && radii, "means/radii are NULL");
    hipLaunchKernelGGL(k_count_live, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                       reinterpret_cast<const float2*>(means), radii,
                       reinterpret_cast<const float4*>(w.packed), n, gx, gy, width, height,
                       w.counts, w.rect, w.live, w.misc);
    CVQPG_LAUNCH_CHECK();
  }
  return scan_exclusive(w.counts, w.offs, w.sums_n, n, w.misc, capacity, num_pairs, nullptr, s);
}

extern "C" int cvqpg_bin_sort(const float* means, const int32_t* radii, int n, int width,
                              int height, void* ws, size_t ws_bytes, int64_t capacity,
                              int32_t* point_list, int32_t* ranges, void* stream) {
  (void)means;
  (void)radii;
  if (int rc = check_common(n, 1, width, height, ws, ws_bytes, capacity)) return rc;
  CVQPG_CHECK_ARG(point_list && ranges, "point_list/ranges are NULL");
  hipStream_t s = (hipStream_t)stream;
  RasterWS w = carve(ws, n, capacity);
  const int gx = ceil_div(width, TILE), gy = ceil_div(height, TILE);
  const int T = gx * gy;
  const int64_t cap = capacity > 0 ? capacity : 1;
  // untouched tiles keep (0, 0); point_list[0] = 0 (the reference's list when no key
  // exists): k_duplicate clears them first, k_clear_ranges when there is no Gaussian
  if (n > 0)
    hipLaunchKernelGGL(k_duplicate, dim3(ceil_div(n, 256)), dim3(256), 0, s, w.counts, w.offs,
                       w.rect, w.live, w.misc, n, gx, w.keys_a, w.vals_a, ranges,
                       2 * (int64_t)T, point_list);
  else
    hipLaunchKernelGGL(k_clear_ranges, dim3(ceil_div(2 * (int64_t)T, 256)), dim3(256), 0, s,
                       ranges, 2 * (int64_t)T, point_list);
  CVQPG_LAUNCH_CHECK();
  const int bits = tile_bits(T);
  const int nb = sort_blocks(cap);
  uint32_t *kin = w.keys_a, *vin = w.vals_a, *kout = w.keys_b, *vout = w.vals_b;
  // as few passes of at most 8 bits as the key needs, with the bits spread evenly over them
  // (13 bits: 7 + 6 rather than 8 + 5): fewer buckets per pass make longer runs of each
  // bucket per block, so the scatter's stores coalesce better
  const int npass = (bits + 7) / 8;
  const int pbits = npass > 0 ? (bits + npass - 1) / npass : 0;
  for (int shift = 0; shift < bits; shift += pbits) {
    const int nbits = bits - shift < pbits ? bits - shift : pbits;
    // the last pass scatters the Gaussian ids straight into point_list
    if (shift + pbits >= bits) vout = reinterpret_cast<uint32_t*>(point_list);
    hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(SORT_THREADS), 0, s, kin, w.misc, shift,
                       (1u << nbits) - 1u, nb, w.hist);
    CVQPG_LAUNCH_CHECK();
    if (int rc = scan_exclusive(w.hist, w.hist, w.sums_h, (int64_t)256 * nb, nullptr, 0, nullptr,
                                nullptr, s))
      return rc;
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(SORT_THREADS), 0, s, kin, vin, kout, vout,
                       w.hist, w.misc, shift, nbits, nb);
    CVQPG_LAUNCH_CHECK();
    uint32_t* tk = kin;
    kin = kout;
    kout = tk;
    uint32_t* tv = vin;
    vin = vout;
    vout = tv;
  }
  if (bits == 0)  // a single tile: nothing to sort
    hipLaunchKernelGGL(k_copy_u32, dim3(ceil_div(cap, 256)), dim3(256), 0, s, vin, point_list,
                       w.misc);
  hipLaunchKernelGGL(k_ranges, dim3(ceil_div(cap, 256)), dim3(256), 0, s, kin, w.misc,
                     reinterpret_cast<int2*>(ranges));
  CVQPG_LAUNCH_CHECK();
  return 0;
}

template <int C>
static void launch_pack(const float* means, const float* inv_cov, const float* colours,
                        const float* phase, const float* opac, const float* curv,
                        const float* area, const float* k, float dxy2, int n, float* packed,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_pack<C>, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                     reinterpret_cast<const float2*>(means), inv_cov, colours, phase, opac, curv,
                     area, k, dxy2, n, packed);
}

extern "C" int cvqpg_raster_pack(const float* means, const float* inv_cov, const float* colours,
                                 const float* phase, const float* opacities,
                                 const float* curvature, const float* area, const float* k, int n,
                                 int channels, float dxy2, int width, int height, void* ws,
                                 size_t ws_bytes, int64_t capacity, void* stream) {
  if (int rc = check_common(n, channels, width, height, ws, ws_bytes, capacity)) return rc;
  if (n == 0) return 0;
  CVQPG_CHECK_ARG(means && inv_cov && colours && phase && opacities && curvature && area && k,
                  "NULL Gaussian parameter pointer");
  hipStream_t s = (hipStream_t)stream;
  RasterWS w = carve(ws, n, capacity);
  if (channels == 3)
    launch_pack<3>(means, inv_cov, colours, phase, opacities, curvature, area, k, dxy2, n, w.packed, s);
  else
    launch_pack<1>(means, inv_cov, colours, phase, opacities, curvature, area, k, dxy2, n, w.packed, s);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_render_forward_packed(int n, int channels, int width, int height, void* ws,
                                           size_t ws_bytes, int64_t capacity,
                                           const int32_t* point_list, const int32_t* ranges,
                                           float* field, void* stream) {
  if (int rc = check_common(n, channels, width, height, ws, ws_bytes, capacity)) return rc;
  CVQPG_CHECK_ARG(field && ranges && point_list, "NULL field/ranges/point_list");
  hipStream_t s = (hipStream_t)stream;
  RasterWS w = carve(ws, n, capacity);
  const int gx = ceil_div(width, TILE), gy = ceil_div(height, TILE);
  if (channels == 3)
    hipLaunchKernelGGL(k_render_fwd<3>, dim3(gx * gy), dim3(256), 0, s,
                       reinterpret_cast<const int2*>(ranges), point_list,
                       reinterpret_cast<const float4*>(w.packed), w.misc, width, height, gx,
                       reinterpret_cast<float2*>(field));
  else
    hipLaunchKernelGGL(k_render_fwd<1>, dim3(gx * gy), dim3(256), 0, s,
                       reinterpret_cast<const int2*>(ranges), point_list,
                       reinterpret_cast<const float4*>(w.packed), w.misc, width, height, gx,
                       reinterpret_cast<float2*>(field));
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_render_backward_tiles(const float* grad_field, const float* k, int n,
                                           int channels, int width, int height, void* ws,
                                           size_t ws_bytes, int64_t capacity,
                                           const int32_t* point_list, const int32_t* ranges,
                                           void* stream) {
  if (int rc = check_common(n, channels, width, height, ws, ws_bytes, capacity)) return rc;
  if (n == 0) return 0;
  CVQPG_CHECK_ARG(grad_field && ranges && point_list && k, "NULL grad/ranges/point_list/k");
  hipStream_t s = (hipStream_t)stream;
  RasterWS w = carve(ws, n, capacity);
  const int gx = ceil_div(width, TILE), gy = ceil_div(height, TILE);
  const int2* rg = reinterpret_cast<const int2*>(ranges);
  const float4* pk = reinterpret_cast<const float4*>(w.packed);
  const float2* gf = reinterpret_cast<const float2*>(grad_field);
  float4* part = reinterpret_cast<float4*>(w.partials);
  if (channels == 3)
    hipLaunchKernelGGL(k_render_bwd<3>, dim3(gx * gy), dim3(256), 0, s, rg, point_list, pk, w.rect,
                       w.live, gf, k, w.misc, width, height, gx, part);
  else
    hipLaunchKernelGGL(k_render_bwd<1>, dim3(gx * gy), dim3(256), 0, s, rg, point_list, pk, w.rect,
                       w.live, gf, k, w.misc, width, height, gx, part);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" size_t cvqpg_raster_batch_table_size(int n_scenes) {
  return sizeof(BatchHeader) + (size_t)(n_scenes > 0 ? n_scenes : 0) * sizeof(BatchScene);
}

extern "C" int cvqpg_raster_batch_table(const cvqpg_raster_scene* scenes, int n_scenes,
                                        void* table_host, size_t table_bytes) {
  CVQPG_CHECK_ARG(n_scenes >= 1 && n_scenes <= MAX_BATCH, "n_scenes must be in [1, %d] (got %d)",
                  MAX_BATCH, n_scenes);
  CVQPG_CHECK_ARG(scenes && table_host, "NULL scenes/table");
  CVQPG_CHECK_ARG(table_bytes >= cvqpg_raster_batch_table_size(n_scenes),
                  "table too small: %zu < %zu bytes", table_bytes,
                  cvqpg_raster_batch_table_size(n_scenes));
  BatchHeader* hdr = static_cast<BatchHeader*>(table_host);
  BatchScene* sc = reinterpret_cast<BatchScene*>(hdr + 1);
  memset(table_host, 0, cvqpg_raster_batch_table_size(n_scenes));
  int64_t tiles = 0;
  for (int i = 0; i < n_scenes; ++i) {
    const cvqpg_raster_scene& q = scenes[i];
    if (int rc = check_common(q.n, q.channels, q.width, q.height, q.ws, q.ws_bytes, q.capacity))
      return rc;
    CVQPG_CHECK_ARG(q.channels == scenes[0].channels, "scene %d has %d channels, scene 0 has %d",
                    i, q.channels, scenes[0].channels);
    CVQPG_CHECK_ARG(q.point_list && q.ranges && q.field, "scene %d: NULL point_list/ranges/field",
                    i);
    RasterWS w = carve(q.ws, q.n, q.capacity);
    const int gx = ceil_div(q.width, TILE), gy = ceil_div(q.height, TILE);
    BatchScene& b = sc[i];
    b.ranges = reinterpret_cast<const int2*>(q.ranges);
    b.point_list = q.point_list;
    b.packed = reinterpret_cast<const float4*>(w.packed);
    b.rect = w.rect;
    b.misc = w.misc;
    b.field = reinterpret_cast<float2*>(q.field);
    b.grad = reinterpret_cast<const float2*>(q.grad_field);
    b.k = q.k;
    b.partials = reinterpret_cast<float4*>(w.partials);
    b.live = w.live;
    b.W = q.width;
    b.H = q.height;
    b.gx = gx;
    b.tile_begin = (int)tiles;
    tiles += (int64_t)gx * gy;
    CVQPG_CHECK_ARG(tiles < ((int64_t)1 << 30), "too many tiles in one batch");
  }
  hdr->n_scenes = n_scenes;
  hdr->total_tiles = (int)tiles;
  return 0;
}

static int check_batch_launch(void* table, int channels, int grid, int chunk) {
  CVQPG_CHECK_ARG(table != nullptr, "table is NULL");
  CVQPG_CHECK_ARG(channels == 1 || channels == 3,
                  "qpf rasterizer supports 1 or 3 channels, got %d", channels);
  CVQPG_CHECK_ARG(grid >= 1 && grid <= 65536, "grid must be in [1, 65536] (got %d)", grid);
  CVQPG_CHECK_ARG(chunk >= 1 && chunk <= 64, "chunk must be in [1, 64] (got %d)", chunk);
  return 0;
}

extern "C" int cvqpg_render_forward_batch(void* table, int channels, int grid, int chunk,
                                          void* stream) {
  if (int rc = check_batch_launch(table, channels, grid, chunk)) return rc;
  hipStream_t s = (hipStream_t)stream;
  BatchHeader* hdr = static_cast<BatchHeader*>(table);
  const BatchScene* sc = reinterpret_cast<const BatchScene*>(hdr + 1);
  if (channels == 3)
    hipLaunchKernelGGL(k_render_fwd_batch<3>, dim3(grid), dim3(256), 0, s, hdr, sc, chunk);
  else
    hipLaunchKernelGGL(k_render_fwd_batch<1>, dim3(grid), dim3(256), 0, s, hdr, sc, chunk);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_render_backward_batch(void* table, int channels, int grid, int chunk,
                                           void* stream) {
  if (int rc = check_batch_launch(table, channels, grid, chunk)) return rc;
  hipStream_t s = (hipStream_t)stream;
  BatchHeader* hdr = static_cast<BatchHeader*>(table);
  const BatchScene* sc = reinterpret_cast<const BatchScene*>(hdr + 1);
  if (channels == 3)
    hipLaunchKernelGGL(k_render_bwd_batch<3>, dim3(grid), dim3(256), 0, s, hdr, sc, chunk);
  else
    hipLaunchKernelGGL(k_render_bwd_batch<1>, dim3(grid), dim3(256), 0, s, hdr, sc, chunk);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_render_backward_reduce(const float* colours, const float* curvature,
                                            const float* area, int n, int channels, float dxy2,
                                            int width, int height, void* ws, size_t ws_bytes,
                                            int64_t capacity, float* d_means, float* d_inv_cov,
                                            float* d_colours, float* d_phase, float* d_opacities,
                                            float* d_curvature, float* d_area, void* stream) {
  if (int rc = check_common(n, channels, width, height, ws, ws_bytes, capacity)) return rc;
  if (n == 0) return 0;
  CVQPG_CHECK_ARG(colours && curvature && area, "NULL colours/curvature/area");
  CVQPG_CHECK_ARG(d_means && d_inv_cov && d_colours && d_phase && d_opacities && d_curvature &&
                      d_area,
                  "NULL gradient output pointer");
  hipStream_t s = (hipStream_t)stream;
  RasterWS w = carve(ws, n, capacity);
  const dim3 gb(ceil_div(n, 256)), tb(256);
  if (channels == 3)
    hipLaunchKernelGGL(k_bwd_reduce<3>, gb, tb, 0, s, reinterpret_cast<const float4*>(w.partials),
                       w.rect, w.counts, w.misc, colours, curvature, area, dxy2, n, d_means,
                       d_inv_cov, d_colours, d_phase, d_opacities, d_curvature, d_area);
  else
    hipLaunchKernelGGL(k_bwd_reduce<1>, gb, tb, 0, s, reinterpret_cast<const float4*>(w.partials),
                       w.rect, w.counts, w.misc, colours, curvature, area, dxy2, n, d_means,
                       d_inv_cov, d_colours, d_phase, d_opacities, d_curvature, d_area);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_render_forward(const float* means, const float* inv_cov,
                                    const float* colours, const float* phase,
                                    const float* opacities, const float* curvature,
                                    const float* area, const float* k, int n, int channels,
                                    float dxy2, int width, int height, void* ws, size_t ws_bytes,
                                    int64_t capacity, const int32_t* point_list,
                                    const int32_t* ranges, float* field, void* stream) {
  if (int rc = cvqpg_raster_pack(means, inv_cov, colours, phase, opacities, curvature, area, k, n,
                                 channels, dxy2, width, height, ws, ws_bytes, capacity, stream))
    return rc;
  return cvqpg_render_forward_packed(n, channels, width, height, ws, ws_bytes, capacity,
                                     point_list, ranges, field, stream);
}

extern "C" int cvqpg_render_backward(const float* grad_field, const float* means,
                                     const float* inv_cov, const float* colours,
                                     const float* phase, const float* opacities,
                                     const float* curvature, const float* area, const float* k,
                                     int n, int channels, float dxy2, int width, int height,
                                     void* ws, size_t ws_bytes, int64_t capacity,
                                     const int32_t* point_list, const int32_t* ranges,
                                     float* d_means, float* d_inv_cov, float* d_colours,
                                     float* d_phase, float* d_opacities, float* d_curvature,
                                     float* d_area, void* stream) {
  if (int rc = cvqpg_raster_pack(means, inv_cov, colours, phase, opacities, curvature, area, k, n,
                                 channels, dxy2, width, height, ws, ws_bytes, capacity, stream))
    return rc;
  if (int rc = cvqpg_render_backward_tiles(grad_field, k, n, channels, width, height, ws,
                                           ws_bytes, capacity, point_list, ranges, stream))
    return rc;
  return cvqpg_render_backward_reduce(colours, curvature, area, n, channels, dxy2, width, height,
                                      ws, ws_bytes, capacity, d_means, d_inv_cov, d_colours,
                                      d_phase, d_opacities, d_curvature, d_area, stream);
}

extern "C" int cvqpg_qpf_render_forward(const float* means, const float* inv_cov,
                                        const float* colours, const float* phase,
                                        const float* opacities, const float* curvature,
                                        const float* area, const int32_t* radii, const float* k,
                                        int n, int channels, float dxy2, int width, int height,
                                        void* ws, size_t ws_bytes, int64_t capacity, float* field,
                                        int32_t* point_list, int32_t* ranges, int32_t* num_pairs,
                                        void* stream) {
  if (int rc = cvqpg_bin_count(means, radii, n, width, height, ws, ws_bytes, capacity, num_pairs,
                               stream))
    return rc;
  if (int rc = cvqpg_bin_sort(means, radii, n, width, height, ws, ws_bytes, capacity, point_list,
                              ranges, stream))
    return rc;
  return cvqpg_render_forward(means, inv_cov, colours, phase, opacities, curvature, area, k, n,
                              channels, dxy2, width, height, ws, ws_bytes, capacity, point_list,
                              ranges, field, stream);
}

static int make_prologue(Prologue& pr, int width, int height, float dxy2, float curv_lambda,
                         int curv_mode, float curv_nyquist) {
  CVQPG_CHECK_ARG(curv_mode >= 0 && curv_mode <= 2, "curv_mode must be 0, 1 or 2");
  pr.W = width;
  pr.H = height;
  pr.dxy2 = dxy2;
  pr.curv_lambda = curv_lambda;
  pr.curv_nyquist = curv_nyquist;
  pr.curv_mode = curv_mode;
  return 0;
}

static RawPtrs raw_ptrs(const float* const* raw) {
  RawPtrs R;
  for (int i = 0; i < NPARAM; ++i) R.p[i] = raw[i];
  return R;
}

extern "C" int cvqpg_prologue_forward(const float* const* raw_params, int n, int channels,
                                      int width, int height, const float* k, float dxy2,
                                      float curv_lambda, int curv_mode, float curv_nyquist,
                                      void* ws, size_t ws_bytes, int64_t capacity,
                                      float* means_out, int32_t* radii_out, void* stream) {
  if (int rc = check_common(n, channels, width, height, ws, ws_bytes, capacity)) return rc;
  if (n == 0) return 0;
  CVQPG_CHECK_ARG(raw_params && k && means_out && radii_out, "NULL argument pointer");
  for (int i = 0; i < NPARAM; ++i) CVQPG_CHECK_ARG(raw_params[i], "raw parameter %d is NULL", i);
  Prologue pr;
  if (int rc = make_prologue(pr, width, height, dxy2, curv_lambda, curv_mode, curv_nyquist)) return rc;
  hipStream_t s = (hipStream_t)stream;
  RasterWS w = carve(ws, n, capacity);
  const RawPtrs R = raw_ptrs(raw_params);
  const LiveOut lo{0, 0, nullptr, nullptr, nullptr, nullptr};
  if (channels == 3)
    hipLaunchKernelGGL((k_prologue<3, false>), dim3(ceil_div(n, 256)), dim3(256), 0, s, R, pr, k,
                       n, means_out, radii_out, w.packed, lo);
  else
    hipLaunchKernelGGL((k_prologue<1, false>), dim3(ceil_div(n, 256)), dim3(256), 0, s, R, pr, k,
                       n, means_out, radii_out, w.packed, lo);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_prologue_forward_live(const float* const* raw_params, int n, int channels,
                                           int width, int height, const float* k, float dxy2,
                                           float curv_lambda, int curv_mode, float curv_nyquist,
                                           void* ws, size_t ws_bytes, int64_t capacity,
                                           float* means_out, int32_t* radii_out, void* stream) {
  if (int rc = check_common(n, channels, width, height, ws, ws_bytes, capacity)) return rc;
  if (n == 0) return 0;
  CVQPG_CHECK_ARG(raw_params && k && means_out && radii_out, "NULL argument pointer");
  for (int i = 0; i < NPARAM; ++i) CVQPG_CHECK_ARG(raw_params[i], "raw parameter %d is NULL", i);
  Prologue pr;
  if (int rc = make_prologue(pr, width, height, dxy2, curv_lambda, curv_mode, curv_nyquist)) return rc;
  hipStream_t s = (hipStream_t)stream;
  RasterWS w = carve(ws, n, capacity);
  const RawPtrs R = raw_ptrs(raw_params);
  const LiveOut lo{ceil_div(width, TILE), ceil_div(height, TILE), w.counts, w.rect, w.live, w.misc};
  if (channels == 3)
    hipLaunchKernelGGL((k_prologue<3, true>), dim3(ceil_div(n, 256)), dim3(256), 0, s, R, pr, k,
                       n, means_out, radii_out, w.packed, lo);
  else
    hipLaunchKernelGGL((k_prologue<1, true>), dim3(ceil_div(n, 256)), dim3(256), 0, s, R, pr, k,
                       n, means_out, radii_out, w.packed, lo);
  CVQPG_LAUNCH_CHECK();
  return 0;
}

extern "C" int cvqpg_bin_scan(int n, void* ws, size_t ws_bytes, int64_t capacity,
                              int32_t* num_pairs, void* stream) {
  if (int rc = check_common(n, 1, 1, 1, ws, ws_bytes, capacity)) return rc;
  RasterWS w = carve(ws, n, capacity);
  return scan_exclusive(w.counts, w.offs, w.sums_n, n, w.misc, capacity, num_pairs, nullptr,
                        (hipStream_t)stream);
}

extern "C" int cvqpg_raster_backward_raw(const float* const* raw_params, int n, int channels,
                                         int width, int height, float dxy2, float curv_lambda,
                                         int curv_mode, float curv_nyquist,
                                         const int32_t* iteration, int curv_warmup, void* ws,
                                         size_t ws_bytes, int64_t capacity, float* grads,
                                         float* sumsq_partials, void* stream) {
  if (int rc = check_common(n, channels, width, height, ws, ws_bytes, capacity)) return rc;
  if (n == 0) return 0;
  CVQPG_CHECK_ARG(raw_params && grads && sumsq_partials, "NULL argument pointer");
  for (int i = 0; i < NPARAM; ++i) CVQPG_CHECK_ARG(raw_params[i], "raw parameter %d is NULL", i);
  Prologue pr;
  if (int rc = make_prologue(pr, width, height, dxy2, curv_lambda, curv_mode, curv_nyquist)) return rc;
  hipStream_t s = (hipStream_t)stream;
  RasterWS w = carve(ws, n, capacity);
  const RawPtrs R = raw_ptrs(raw_params);
  if (channels == 3)
    hipLaunchKernelGGL(k_bwd_reduce_raw<3>, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                       reinterpret_cast<const float4*>(w.partials), w.rect, w.counts, w.misc, R,
                       pr, iteration, curv_warmup, n, grads, sumsq_partials);
  else
    hipLaunchKernelGGL(k_bwd_reduce_raw<1>, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                       reinterpret_cast<const float4*>(w.partials), w.rect, w.counts, w.misc, R,
                       pr, iteration, curv_warmup, n, grads, sumsq_partials);
  CVQPG_LAUNCH_CHECK();
  return 0;
}
