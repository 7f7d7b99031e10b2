// Mixed-radix Stockham FFT that runs inside one workgroup on data held in LDS.
//
// A sequence of length n (or CG interleaved sequences: element i of sequence g is at
// buf[i * CG + g]) is transformed in place.  Each radix-R stage with
// ns = product of the earlier radices does, for butterfly j in [0, n/R):
//     k = j mod ns
//     v[r] = x[j + r n/R] * W_{ns R}^{r k}            (r = 0..R-1)
//     v    = DFT_R(v)
//     y[(j / ns) ns R + k + r ns] = v[r]
// which leaves the result in natural order (Stockham autosort).  All of a thread's
// inputs for a stage are read into registers before the barrier and written after it,
// so a single LDS buffer suffices.
//
// Twiddles come from a two-level table (W_n^{64 h} and W_n^l, l < 64) evaluated in FP64,
// rounded to FP32 once and staged in LDS by the caller; a butterfly forms W^{r k} as
// powers of W^k.  DIR = +1 is the forward transform (e^{-i}), DIR = -1 the
// unnormalised inverse (e^{+i}).  Radices 2, 3, 4, 5 and 8 have dedicated butterflies;
// the odd primes 7, 11, 13 and 17 use a symmetric direct DFT with compile-time roots.
//
// A length n with a prime factor above 17 runs as Bluestein's chirp-z transform on a
// Stockham transform of 7-smooth length m >= 2n - 1 held in the same LDS buffer.  With
// c_j = e^{-i pi j^2 / n} (j^2 reduced modulo 2n exactly) and jk = (j^2 + k^2 - (k-j)^2)/2,
//     X_k = sum_j x_j e^{-2 pi i jk/n} = c_k sum_j (x_j c_j) conj(c_{k-j}),
// a linear convolution with the even sequence b_j = conj(c_j), |j| < n, which the m-point
// circular convolution reproduces exactly because m >= 2n - 1:
//     a = (x c, 0 ...),  X = c * IFFT_m(FFT_m(a) * B),  B = FFT_m(b) / m.
// The unnormalised inverse conjugates c and B (B is even, so FFT_m(conj b) = conj(B)).
// c (n entries) and B (m entries) are tables in global memory built in FP64 per length.
#pragma once
#include <hip/hip_runtime.h>

#include <tuple>

namespace cvqpg {
namespace fft {

constexpr int MAX_STAGES = 16;

struct Plan {
  int n;             // transform length
  int len;           // Stockham length: n, or the Bluestein length m (chirp != nullptr)
  int nstages;
  int radix[MAX_STAGES];
  int nh;            // ceil(len / 64)
  const float2* tw;  // two-level table of len: tw[h] = W^{64 h} (h < nh), tw[nh + l] = W^l (l < 64)
  const float2* chirp;  // Bluestein only: c_j, j < n
  const float2* bhat;   // Bluestein only: B_k, k < m
};

constexpr int TW_LO = 64;
__host__ __device__ constexpr int twiddle_entries(int n) { return (n + TW_LO - 1) / TW_LO + TW_LO; }

// LDS element index with one padding slot every 16 complex elements, so the strided
// accesses of the early Stockham stages spread over the banks.
__device__ __forceinline__ int lidx(int i) { return i + (i >> 4); }
__host__ __device__ inline int lds_elems(int n) { return n + (n >> 4) + 1; }

// Factorisation of a Stockham length into supported radices; powers of two become
// radix-8 and radix-4 stages (never radix 2 unless n has a single factor 2); returns
// false if n has a prime factor above 17.
inline bool make_stages(int n, Plan& p) {
  p.n = n;
  p.len = n;
  p.nstages = 0;
  p.nh = (n + TW_LO - 1) / TW_LO;
  p.chirp = nullptr;
  p.bhat = nullptr;
  if (n < 2) return false;
  int m = n, a = 0;
  while (m % 2 == 0) {
    m /= 2;
    ++a;
  }
  int n8 = a / 3, n4 = 0, n2 = 0;
  if (a % 3 == 1) {
    if (n8 > 0) {
      --n8;
      n4 = 2;
    } else {
      n2 = 1;
    }
  } else if (a % 3 == 2) {
    n4 = 1;
  }
  for (int i = 0; i < n8; ++i) p.radix[p.nstages++] = 8;
  for (int i = 0; i < n4; ++i) p.radix[p.nstages++] = 4;
  for (int i = 0; i < n2; ++i) p.radix[p.nstages++] = 2;
  const int odd[] = {3, 5, 7, 11, 13, 17};
  for (int r : odd) {
    while (m % r == 0) {
      if (p.nstages >= MAX_STAGES) return false;
      p.radix[p.nstages++] = r;
      m /= r;
    }
  }
  return m == 1;
}

// smallest 2^a 3^b 5^c 7^d >= v
inline int smooth7_at_least(int v) {
  int best = 0;
  for (long a = 1; a < 2L * v + 2; a *= 2)
    for (long b = a; b < 2L * v + 2; b *= 3)
      for (long c = b; c < 2L * v + 2; c *= 5)
        for (long d = c; d < 2L * v + 2; d *= 7)
          if (d >= v && (best == 0 || d < best)) best = (int)d;
  return best;
}

// Stockham length of a transform of length n: n itself when it factors into the
// supported radices, else the Bluestein length (smallest 7-smooth m >= 2n - 1).
inline int stockham_len(int n) {
  Plan p;
  return make_stages(n, p) ? n : smooth7_at_least(2 * n - 1);
}

// Plan of a transform of length n >= 2 (Bluestein when n has a prime factor above 17; the
// caller sets tw, and for Bluestein chirp and bhat, to tables of p.len / p.n).
inline bool make_plan(int n, Plan& p) {
  if (make_stages(n, p)) return true;
  if (n < 2) return false;
  if (!make_stages(smooth7_at_least(2 * n - 1), p)) return false;
  p.n = n;
  return true;
}

inline bool is_bluestein(const Plan& p) { return p.len != p.n; }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// multiply by -i (DIR=+1) or +i (DIR=-1)
template <int DIR>
__device__ __forceinline__ float2 rot90(float2 a) {
  return DIR > 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// W_n^m from the two-level table (one complex product of two FP32-rounded roots)
template <int DIR>
__device__ __forceinline__ float2 twid(const float2* tw, int nh, int m) {
  const float2 a = tw[m >> 6], b = tw[nh + (m & (TW_LO - 1))];
  const float2 w = make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
  return DIR > 0 ? w : make_float2(w.x, -w.y);
}

template <int DIR>
__device__ __forceinline__ void bfly2(float2& a, float2& b) {
  const float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2* v) {
  const float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
  const float2 b0 = cadd(v[1], v[3]), b1 = rot90<DIR>(csub(v[1], v[3]));
  v[0] = cadd(a0, b0);
  v[2] = csub(a0, b0);
  v[1] = cadd(a1, b1);
  v[3] = csub(a1, b1);
}

template <int DIR>
__device__ __forceinline__ void dft8(float2* v) {
  constexpr float S = 0.70710678118654752f;
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4<DIR>(e);
  dft4<DIR>(o);
  // o[k] *= W_8^k
  o[1] = DIR > 0 ? make_float2(S * (o[1].x + o[1].y), S * (o[1].y - o[1].x))
                 : make_float2(S * (o[1].x - o[1].y), S * (o[1].y + o[1].x));
  o[2] = rot90<DIR>(o[2]);
  o[3] = DIR > 0 ? make_float2(S * (-o[3].x + o[3].y), S * (-o[3].y - o[3].x))
                 : make_float2(S * (-o[3].x - o[3].y), S * (-o[3].y + o[3].x));
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 4] = csub(e[k], o[k]);
  }
}

template <int DIR>
__device__ __forceinline__ void dft16(float2* v) {
  // 16 = 4 x 4: DFT4 over n1 for each n2 (x[4 n1 + n2]), twiddle W16^{n2 k1}, DFT4 over n2
  constexpr float C1 = 0.92387953251128676f, S1 = 0.38268343236508977f;
  constexpr float S2 = 0.70710678118654752f;
  float2 a[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    a[n2][0] = v[n2];
    a[n2][1] = v[4 + n2];
    a[n2][2] = v[8 + n2];
    a[n2][3] = v[12 + n2];
    dft4<DIR>(a[n2]);
  }
  // W16^m for m = n2 k1 in {1, 2, 3, 4, 6, 9}; forward uses e^{-i}
  auto tw = [](float2 x, float c, float s) {
    // x * (c - i s) forward, x * (c + i s) inverse
    return DIR > 0 ? make_float2(x.x * c + x.y * s, x.y * c - x.x * s)
                   : make_float2(x.x * c - x.y * s, x.y * c + x.x * s);
  };
  a[1][1] = tw(a[1][1], C1, S1);
  a[1][2] = tw(a[1][2], S2, S2);
  a[1][3] = tw(a[1][3], S1, C1);
  a[2][1] = tw(a[2][1], S2, S2);
  a[2][2] = rot90<DIR>(a[2][2]);
  a[2][3] = tw(a[2][3], -S2, S2);
  a[3][1] = tw(a[3][1], S1, C1);
  a[3][2] = tw(a[3][2], -S2, S2);
  a[3][3] = tw(a[3][3], -C1, -S1);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 b[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
    dft4<DIR>(b);
    v[k1] = b[0];
    v[k1 + 4] = b[1];
    v[k1 + 8] = b[2];
    v[k1 + 12] = b[3];
  }
}

template <int DIR>
__device__ __forceinline__ void dft3(float2* v) {
  constexpr float C1 = -0.5f;
  constexpr float S1 = 0.86602540378443865f;  // sin(2 pi / 3)
  const float2 s = cadd(v[1], v[2]);
  const float2 d = csub(v[1], v[2]);
  const float2 m = make_float2(v[0].x + C1 * s.x, v[0].y + C1 * s.y);
  // forward: v1 = m - i S1 d, v2 = m + i S1 d
  const float2 t = DIR > 0 ? make_float2(S1 * d.y, -S1 * d.x) : make_float2(-S1 * d.y, S1 * d.x);
  v[0] = cadd(v[0], s);
  v[1] = cadd(m, t);
  v[2] = csub(m, t);
}

template <int DIR>
__device__ __forceinline__ void dft5(float2* v) {
  constexpr float C1 = 0.30901699437494742f;   // cos(2 pi / 5)
  constexpr float C2 = -0.80901699437494742f;  // cos(4 pi / 5)
  constexpr float S1 = 0.95105651629515357f;   // sin(2 pi / 5)
  constexpr float S2 = 0.58778525229247313f;   // sin(4 pi / 5)
  const float2 s1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
  const float2 s2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
  const float2 a1 = make_float2(v[0].x + C1 * s1.x + C2 * s2.x, v[0].y + C1 * s1.y + C2 * s2.y);
  const float2 a2 = make_float2(v[0].x + C2 * s1.x + C1 * s2.x, v[0].y + C2 * s1.y + C1 * s2.y);
  // b = S1 d1 + S2 d2 ; c = S2 d1 - S1 d2 ; forward multiplies them by -i
  const float2 b = make_float2(S1 * d1.x + S2 * d2.x, S1 * d1.y + S2 * d2.y);
  const float2 c = make_float2(S2 * d1.x - S1 * d2.x, S2 * d1.y - S1 * d2.y);
  const float2 bi = rot90<DIR>(b), ci = rot90<DIR>(c);
  v[0] = cadd(v[0], cadd(s1, s2));
  v[1] = cadd(a1, bi);
  v[4] = csub(a1, bi);
  v[2] = cadd(a2, ci);
  v[3] = csub(a2, ci);
}

// Compile-time R-th roots of unity (Taylor series in double after reduction to [-pi, pi]),
// so the direct DFT of the odd radices uses literal constants instead of registers.
constexpr double kPi = 3.14159265358979323846;
constexpr double cx_sin(double x) {
  double term = x, sum = x;
  for (int i = 1; i < 40; ++i) {
    term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cx_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 40; ++i) {
    term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
    sum += term;
  }
  return sum;
}
template <int R>
struct Roots {
  float c[R];
  float s[R];
  constexpr Roots() : c(), s() {
    for (int m = 0; m < R; ++m) {
      double a = 2.0 * kPi * (double)m / (double)R;
      if (a > kPi) a -= 2.0 * kPi;
      c[m] = (float)cx_cos(a);
      s[m] = (float)cx_sin(a);
    }
  }
};

// direct DFT of odd length R in the symmetric form
//   X_k, X_{R-k} = v0 + sum_j s_j cos(2 pi j k / R)  -/+ i sum_j d_j sin(2 pi j k / R)
// with s_j = v_j + v_{R-j}, d_j = v_j - v_{R-j} (forward; the inverse swaps the signs)
template <int R, int DIR>
__device__ __forceinline__ void dft_odd(float2* v) {
  constexpr int h = (R - 1) / 2;
  constexpr Roots<R> rt{};
  float2 sj[h], dj[h];
#pragma unroll
  for (int j = 1; j <= h; ++j) {
    sj[j - 1] = cadd(v[j], v[R - j]);
    dj[j - 1] = csub(v[j], v[R - j]);
  }
  const float2 v0 = v[0];
  float2 tot = v0;
#pragma unroll
  for (int j = 0; j < h; ++j) tot = cadd(tot, sj[j]);
#pragma unroll
  for (int k = 1; k <= h; ++k) {
    float2 a = v0, b = make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 1; j <= h; ++j) {
      const float c = rt.c[(j * k) % R], sn = rt.s[(j * k) % R];
      a.x = fmaf(sj[j - 1].x, c, a.x);
      a.y = fmaf(sj[j - 1].y, c, a.y);
      b.x = fmaf(dj[j - 1].x, sn, b.x);
      b.y = fmaf(dj[j - 1].y, sn, b.y);
    }
    // forward: X_k = a - i b ; X_{R-k} = a + i b
    const float2 ib = make_float2(-b.y, b.x);
    v[k] = DIR > 0 ? csub(a, ib) : cadd(a, ib);
    v[R - k] = DIR > 0 ? cadd(a, ib) : csub(a, ib);
  }
  v[0] = tot;
}

template <int R, int DIR>
__device__ __forceinline__ void dft(float2* v) {
  if constexpr (R == 2) {
    bfly2<DIR>(v[0], v[1]);
  } else if constexpr (R == 3) {
    dft3<DIR>(v);
  } else if constexpr (R == 4) {
    dft4<DIR>(v);
  } else if constexpr (R == 5) {
    dft5<DIR>(v);
  } else if constexpr (R == 8) {
    dft8<DIR>(v);
  } else if constexpr (R == 16) {
    dft16<DIR>(v);
  } else {
    dft_odd<R, DIR>(v);
  }
}

// One radix-R stage on CG interleaved sequences of length n held in buf (padded index).
// LCAP is the LDS capacity in complex elements (n * CG <= LCAP); NT threads.  The twiddles
// W^{r k} of a butterfly are built as powers of W^k.
template <int R, int DIR, int CG, int NT, int LCAP>
__device__ __forceinline__ void stage(float2* buf, int n, int ns, const float2* tw, int nh) {
  constexpr int MAXB = (LCAP / R + NT - 1) / NT;
  const int cnt = n / R;
  const int total = cnt * CG;
  const int tstep = n / (ns * R);
  float2 v[MAXB][R];
  int dst[MAXB];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int w = threadIdx.x + b * NT;
    dst[b] = -1;
    if (w < total) {
      const int col = w % CG;
      const int j = w / CG;
      const int k = j % ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = buf[lidx((j + r * cnt) * CG + col)];
      if (ns > 1) {
        const float2 w1 = twid<DIR>(tw, nh, k * tstep);
        float2 wr = w1;
        v[b][1] = cmul(v[b][1], w1);
#pragma unroll
        for (int r = 2; r < R; ++r) {
          wr = cmul(wr, w1);
          v[b][r] = cmul(v[b][r], wr);
        }
      }
      dft<R, DIR>(v[b]);
      dst[b] = ((j / ns) * ns * R + k) * CG + col;
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    if (dst[b] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[lidx(dst[b] + r * ns * CG)] = v[b][r];
    }
  }
  __syncthreads();
}

// Full transform; the caller has written the input to buf, staged the twiddle table at tw
// (LDS) and synchronised.  MAXR bounds
// the radices compiled into this instance: kernels whose plans only use radices <= 8
// instantiate MAXR = 8 and do not pay the registers of the large direct-DFT stages.
template <int DIR, int CG, int NT, int LCAP, int MAXR = 17>
__device__ __forceinline__ void run(float2* buf, const Plan& p, const float2* tw) {
  int ns = 1;
  const int n = p.len;
  for (int s = 0; s < p.nstages; ++s) {
    const int R = p.radix[s];
    switch (R) {
      case 2: stage<2, DIR, CG, NT, LCAP>(buf, n, ns, tw, p.nh); break;
      case 3: stage<3, DIR, CG, NT, LCAP>(buf, n, ns, tw, p.nh); break;
      case 4: stage<4, DIR, CG, NT, LCAP>(buf, n, ns, tw, p.nh); break;
      case 5: stage<5, DIR, CG, NT, LCAP>(buf, n, ns, tw, p.nh); break;
      case 7: stage<7, DIR, CG, NT, LCAP>(buf, n, ns, tw, p.nh); break;
      case 8: stage<8, DIR, CG, NT, LCAP>(buf, n, ns, tw, p.nh); break;
      case 11:
        if constexpr (MAXR >= 11) stage<11, DIR, CG, NT, LCAP>(buf, n, ns, tw, p.nh);
        break;
      case 13:
        if constexpr (MAXR >= 13) stage<13, DIR, CG, NT, LCAP>(buf, n, ns, tw, p.nh);
        break;
      case 17:
        if constexpr (MAXR >= 17) stage<17, DIR, CG, NT, LCAP>(buf, n, ns, tw, p.nh);
        break;
      default: break;
    }
    ns *= R;
  }
}

// Bluestein transform of CG interleaved sequences of length p.n whose inputs occupy the
// first p.n positions of buf; the Stockham stages run on p.len positions (the caller's
// buffer holds p.len * CG elements).  Same contract as run: synchronised on entry, the
// result in the first p.n positions and synchronised on exit.
template <int DIR, int CG, int NT, int LCAP, int MAXR>
__device__ __forceinline__ void bluestein(float2* buf, const Plan& p, const float2* tw) {
  const int n = p.n, total = p.len * CG;
  for (int e = threadIdx.x; e < total; e += NT) {
    const int i = e / CG;
    float2 v = make_float2(0.f, 0.f);
    if (i < n) {
      const float2 c = p.chirp[i];
      v = cmul(buf[lidx(e)], DIR > 0 ? c : make_float2(c.x, -c.y));
    }
    buf[lidx(e)] = v;
  }
  __syncthreads();
  run<1, CG, NT, LCAP, MAXR>(buf, p, tw);
  for (int e = threadIdx.x; e < total; e += NT) {
    const float2 b = p.bhat[e / CG];
    buf[lidx(e)] = cmul(buf[lidx(e)], DIR > 0 ? b : make_float2(b.x, -b.y));
  }
  __syncthreads();
  run<-1, CG, NT, LCAP, MAXR>(buf, p, tw);
  for (int e = threadIdx.x; e < n * CG; e += NT) {
    const float2 c = p.chirp[e / CG];
    buf[lidx(e)] = cmul(buf[lidx(e)], DIR > 0 ? c : make_float2(c.x, -c.y));
  }
  __syncthreads();
}

inline int max_radix(const Plan& p) {
  int m = 0;
  for (int s = 0; s < p.nstages; ++s) m = p.radix[s] > m ? p.radix[s] : m;
  return m;
}

// ------------------------------------------------------------------------------
// Compile-time plans: the stage sequence, the stride ns and every per-thread butterfly
// count are constants, so the index arithmetic folds to shifts/multiplies and each stage
// holds exactly the registers it needs.
// ------------------------------------------------------------------------------
// ---------------------------------------------------------------------------------
// compile-time transforms
// ---------------------------------------------------------------------------------
// Geo<N, R, CG, NT>: work split of one radix-R stage of length N over CG interleaved
// sequences and NT threads; butterfly b of a thread is w = threadIdx.x + b NT, which
// covers column w mod CG and butterfly j = w / CG.
template <int N, int R, int CG, int NT>
struct Geo {
  static constexpr int CNT = N / R;
  static constexpr int TOTAL = CNT * CG;
  static constexpr int MAXB = (TOTAL + NT - 1) / NT;
  __device__ __forceinline__ static bool live(int b) {
    return TOTAL % NT == 0 || (int)threadIdx.x + b * NT < TOTAL;
  }
  __device__ __forceinline__ static void item(int b, int& j, int& col) {
    const int w = threadIdx.x + b * NT;
    col = w % CG;
    j = w / CG;
  }
};

// LDS layout of one inter-stage buffer: Q padding slots after every P elements (P a
// power of two; P = 0: unpadded).  The LDS has 64 banks of 4 bytes, so a wave's 64
// eight-byte accesses are conflict-free when they cover every bank exactly twice.
// Contiguous accesses want no padding, the strided writes of the early Stockham stages
// want some; each buffer of a compile-time pipe gets the layout that suits its writer
// and its reader.
template <int P, int Q>
struct Lay {
  static_assert(P == 0 || (P & (P - 1)) == 0, "P must be a power of two");
  static constexpr int P_ = P, Q_ = Q;
  __device__ __forceinline__ static int at(int i) {
    if constexpr (P == 0) return i;
    else return i + (i / P) * Q;
  }
  // slot of element base + r STRIDE: when STRIDE is a multiple of P the padding of every
  // element follows from the base's, so the accesses of a butterfly share one address
  // register and differ by immediate offsets
  template <int STRIDE>
  __device__ __forceinline__ static int slot(int base, int r) {
    if constexpr (P == 0) return base + r * STRIDE;
    else if constexpr (STRIDE % P == 0) return at(base) + r * (STRIDE + STRIDE / P * Q);
    else return at(base + r * STRIDE);
  }
  static constexpr int elems(int n) { return P == 0 ? n : n + (n / P) * Q + Q; }
};

// Inputs of stage (R, NS) from LDS buffer layout LI, twiddled and transformed, into
// registers.  For the first stage (NS = 1) and the last (NS = N / R) of a transform,
// element r of butterfly j is sequence position j + r N / R.  Twiddles come from the
// stage table tt (TAB) or from the two-level table tw as powers of W^k.  Stage (R, NS)
// with NS > 1 stores W_{NS R}^{r k} at (r - 1) NS + k, so the lanes of a wave
// (consecutive k) read consecutive words.
template <int R, int NS, int N, int DIR, int CG, int NT, bool TAB, class LI>
__device__ __forceinline__ void sread(const float2* buf, const float2* tw, int nh,
                                      const float2* tt, float2 (&v)[Geo<N, R, CG, NT>::MAXB][R]) {
  using G = Geo<N, R, CG, NT>;
  constexpr int TSTEP = N / (NS * R);
#pragma unroll
  for (int b = 0; b < G::MAXB; ++b) {
    if (!G::live(b)) continue;
    int j, col;
    G::item(b, j, col);
    const int base = j * CG + col;
#pragma unroll
    for (int r = 0; r < R; ++r) v[b][r] = buf[LI::template slot<G::CNT * CG>(base, r)];
    if constexpr (NS > 1) {
      const int k = j % NS;
      if constexpr (TAB) {
        const float2* t = tt + k;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const float2 w = t[(r - 1) * NS];
          v[b][r] = cmul(v[b][r], DIR > 0 ? w : make_float2(w.x, -w.y));
        }
      } else {
        const float2 w1 = twid<DIR>(tw, nh, k * TSTEP);
        float2 wr = w1;
        v[b][1] = cmul(v[b][1], w1);
#pragma unroll
        for (int r = 2; r < R; ++r) {
          wr = cmul(wr, w1);
          v[b][r] = cmul(v[b][r], wr);
        }
      }
    }
    dft<R, DIR>(v[b]);
  }
}

// Outputs of stage (R, NS) from registers to their Stockham positions in LDS buffer
// layout LO.
template <int R, int NS, int N, int CG, int NT, class LO>
__device__ __forceinline__ void swrite(float2* buf, const float2 (&v)[Geo<N, R, CG, NT>::MAXB][R]) {
  using G = Geo<N, R, CG, NT>;
#pragma unroll
  for (int b = 0; b < G::MAXB; ++b) {
    if (!G::live(b)) continue;
    int j, col;
    G::item(b, j, col);
    if constexpr (NS == 1 && LO::P_ > 0 && (R * CG) % LO::P_ == 0 && LO::P_ % CG == 0) {
      // j R CG is a multiple of P and col + r CG stays in the P-block of r CG
      const int a = LO::at(j * R * CG) + col;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[a + r * CG + (r * CG / LO::P_) * LO::Q_] = v[b][r];
    } else {
      const int dst = ((j / NS) * NS * R + j % NS) * CG + col;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[LO::template slot<NS * CG>(dst, r)] = v[b][r];
    }
  }
}

// Workgroup barrier for LDS exchanges only: the wave waits for its own LDS operations
// (lgkmcnt 0) and not for its global memory operations, so global loads and stores issued
// earlier (a prefetch into LDS, the previous plane's stores) stay in flight across it.
// s_waitcnt immediate (gfx9 layout): vmcnt 63 (bits 3:0 and 15:14), expcnt 7 (6:4),
// lgkmcnt 0 (11:8).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_s_waitcnt(0xC07F);
  __builtin_amdgcn_s_barrier();
}

// One LDS-to-LDS stage; all reads finish before the barrier, all writes before the next.
template <int R, int NS, int N, int DIR, int CG, int NT, bool TAB, class LI, class LO>
__device__ __forceinline__ void cstage(float2* buf, const float2* tw, int nh, const float2* tt) {
  float2 v[Geo<N, R, CG, NT>::MAXB][R];
  sread<R, NS, N, DIR, CG, NT, TAB, LI>(buf, tw, nh, tt, v);
  lds_barrier();
  swrite<R, NS, N, CG, NT, LO>(buf, v);
  lds_barrier();
}

template <int... Rs>
struct Rad {};
template <class... Ls>
struct Lays {};

// Pipe<N, Rad<R1, ..., RL>, Lays<B1, ..., B(L-1)>>: a transform of length N = prod(R)
// whose first stage takes its inputs from registers (NS = 1, no twiddles) and whose last
// stage leaves its outputs in registers, both indexed as position j + r N / R.  Buffer
// Bs holds the output of stage s.  With TAB the middle and last stages read their
// twiddles from the concatenated stage tables at tt (TABLE entries, stage order).
template <int N, class RL, class LL>
struct Pipe;
template <int N, int... Rs, class... Ls>
struct Pipe<N, Rad<Rs...>, Lays<Ls...>> {
  static constexpr int L = sizeof...(Rs);
  static_assert(L >= 2, "a pipe has at least two stages");
  static_assert(sizeof...(Ls) == L - 1, "one buffer layout per stage boundary");
  static constexpr int RR[L] = {Rs...};
  static constexpr int prod(int s) {
    int p = 1;
    for (int i = 0; i < s; ++i) p *= RR[i];
    return p;
  }
  static_assert(prod(L) == N, "radices must multiply to N");
  static constexpr int toff(int s) {
    int o = 0;
    for (int i = 0; i < s; ++i) o += prod(i) > 1 ? prod(i) * (RR[i] - 1) : 0;
    return o;
  }
  static constexpr int n_ = N;
  static constexpr int FIRST = RR[0];
  static constexpr int LAST = RR[L - 1];
  static constexpr int TABLE = toff(L);
  template <int I>
  using LayOf = typename std::tuple_element<I, std::tuple<Ls...>>::type;
  template <int CG, int NT>
  using G1 = Geo<N, FIRST, CG, NT>;
  template <int CG, int NT>
  using GL = Geo<N, LAST, CG, NT>;
  // LDS elements the buffers of a transform of CG interleaved sequences need
  static constexpr int lds_elems(int cg) {
    int m = 0;
    for (int e : {Ls::elems(N * cg)...}) m = e > m ? e : m;
    return m;
  }
  // first-stage butterflies in v (inputs) -> LDS; caller synchronises afterwards
  template <int DIR, int CG, int NT>
  __device__ __forceinline__ static void first(float2* buf, float2 (&v)[G1<CG, NT>::MAXB][FIRST]) {
#pragma unroll
    for (int b = 0; b < G1<CG, NT>::MAXB; ++b)
      if (G1<CG, NT>::live(b)) dft<FIRST, DIR>(v[b]);
    swrite<FIRST, 1, N, CG, NT, LayOf<0>>(buf, v);
  }
  template <int S, int DIR, int CG, int NT, bool TAB>
  __device__ __forceinline__ static void mid_from(float2* buf, const float2* tw, int nh,
                                                  const float2* tt) {
    if constexpr (S < L - 1) {
      cstage<RR[S], prod(S), N, DIR, CG, NT, TAB, LayOf<S - 1>, LayOf<S>>(
          buf, tw, nh, TAB ? tt + toff(S) : tt);
      mid_from<S + 1, DIR, CG, NT, TAB>(buf, tw, nh, tt);
    }
  }
  // middle stages (each ends with a barrier)
  template <int DIR, int CG, int NT, bool TAB = false>
  __device__ __forceinline__ static void middle(float2* buf, const float2* tw, int nh,
                                                const float2* tt = nullptr) {
    mid_from<1, DIR, CG, NT, TAB>(buf, tw, nh, tt);
  }
  // last stage -> registers (no barrier)
  template <int DIR, int CG, int NT, bool TAB = false>
  __device__ __forceinline__ static void last(const float2* buf, const float2* tw, int nh,
                                              float2 (&v)[GL<CG, NT>::MAXB][LAST],
                                              const float2* tt = nullptr) {
    sread<LAST, prod(L - 1), N, DIR, CG, NT, TAB, LayOf<L - 2>>(
        buf, tw, nh, TAB ? tt + toff(L - 1) : tt, v);
  }
  // entry idx (< TABLE) of the stage tables, from FP64 roots
  template <int S = 1>
  __device__ __forceinline__ static void table_entry(float2* out, int idx) {
    if constexpr (S < L) {
      constexpr int NS = prod(S), R = RR[S], OFF = toff(S), E = NS * (R - 1);
      if (idx >= OFF && idx < OFF + E) {
        const int r = (idx - OFF) / NS + 1, k = (idx - OFF) % NS;
        const long m = ((long)r * k) % (NS * R);
        double sn, cs;
        sincospi(2.0 * (double)m / (double)(NS * R), &sn, &cs);
        out[idx] = make_float2((float)cs, (float)-sn);
        return;
      }
      table_entry<S + 1>(out, idx);
    }
  }
};

// Runtime<MAXR, LCAP>: any length whose Stockham length is at most LCAP / CG, stage list
// read from the plan; Bluestein for lengths with a prime factor above 17.  A Bluestein
// length is 7-smooth, so only the MAXR = 8 instances (which the host picks for it by the
// plan's largest radix) carry the Bluestein code.
template <int MAXR, int LCAP>
struct Runtime {
  static constexpr int n = 0;
  template <int DIR, int CG, int NT>
  __device__ __forceinline__ static void exec(float2* buf, const Plan& p, const float2* tw) {
    if constexpr (MAXR <= 8) {
      if (p.chirp) {
        bluestein<DIR, CG, NT, LCAP, MAXR>(buf, p, tw);
        return;
      }
    }
    run<DIR, CG, NT, LCAP, MAXR>(buf, p, tw);
  }
};

}  // namespace fft
}  // namespace cvqpg
