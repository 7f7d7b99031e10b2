// Fused optimiser tail for gfx950: nan_to_num (done where the gradients are written),
// global-norm clip and the Adan update of all seven parameter tensors in one launch.
//
// Reference behaviour:
//   nan_to_num + clip_grad_norm_(1.0)   train_2d_gaussian_rgb.py:205-209
//   Adan.step / _single_tensor_adan     utils/optimizer.py:105-262 (betas (0.98, 0.92,
//                                       0.99), eps 1e-8, weight decay 0, no_prox False,
//                                       its own max_grad_norm 0 -> factor 1)
//   learning-rate groups                train_2d_gaussian_rgb.py:73-86
//   MeansCosineScheduler                train_2d_gaussian_rgb.py:58-70
//   curvature warm-up (frozen, no grad) train_2d_gaussian_rgb.py:150-160, 181-185
//
// The iteration index lives in device memory so a captured HIP graph of the whole step
// replays correctly: the learning rate of the means group, the bias corrections and the
// curvature warm-up are all derived from it inside the kernel, and k_advance bumps it
// at the end of the step.
//
// A step whose keys overflowed the raster's capacity (the flag cvqpg_raster_overflow_flags
// points at) rendered nothing: with that flag passed as skip_flag, the update and the
// iteration advance are both skipped, so a replayed graph never applies an update computed
// from an empty field.
#include "common.h"
#include "cvqpg.h"

namespace cvqpg {

struct AdanParams {
  float* param[7];
  int64_t off[8];       // segment offsets into the flat gradient / state arrays
  float lr[7];          // per-segment learning rate (segment 0 uses the cosine schedule)
  float lr0, eta_min;   // means schedule
  int num_itrs;
  float b1, b2, b3, eps;
  float max_norm;
  int curv_seg;         // segment index of the curvature tensor
  int curv_warmup;      // iterations with the curvature frozen
  int curv_enabled;     // 0: curvature never trained (flat primitive)
};

__global__ __launch_bounds__(256) void k_adan(AdanParams A, const float* __restrict__ grads,
                                              float* __restrict__ state,
                                              const float* __restrict__ sumsq, int n_partials,
                                              const int32_t* __restrict__ iteration,
                                              const uint32_t* __restrict__ skip) {
  __shared__ double red[256];
  __shared__ float s_coef;
  if (skip && skip[0]) return;  // uniform over the grid
  const int t = threadIdx.x;
  double acc = 0.0;
  for (int i = t; i < n_partials; i += 256) acc += (double)sumsq[i];
  red[t] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  if (t == 0) {
    const double norm = sqrt(red[0]);
    const double coef = (double)A.max_norm / (norm + 1e-6);
    s_coef = (float)(coef < 1.0 ? coef : 1.0);
  }
  __syncthreads();
  const int64_t total = A.off[7];
  const int itr = iteration[0];
  const int step = itr + 1;  // every group's step counter advances on every call
  const float bc1 = 1.0f - powf(A.b1, (float)step);
  const float bc2 = 1.0f - powf(A.b2, (float)step);
  const float bc3 = 1.0f - powf(A.b3, (float)step);
  float* m = state;
  float* v = state + total;
  float* dd = state + 2 * total;
  float* npg = state + 3 * total;
  // cosine schedule of the means group (MeansCosineScheduler, train_2d_gaussian_rgb.py:58-70)
  const double T = A.num_itrs > 1 ? A.num_itrs : 1;
  const float lr_means =
      (float)(A.eta_min + (A.lr0 - A.eta_min) * (1.0 + cos(3.141592653589793 * itr / T)) / 2.0);
  // grid-stride in batches of ADAN_BATCH elements per thread (element k of a batch at
  // i0 + k * stride): every load of a batch is issued before any of its arithmetic, so a
  // thread waits out one memory latency per batch instead of one per element; every
  // workgroup repeats the (small) norm reduction above once
  constexpr int ADAN_BATCH = 8;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const float sbc3 = sqrtf(bc3);
  for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + t; i0 < total; i0 += ADAN_BATCH * stride) {
    float g[ADAN_BATCH], pg[ADAN_BATCH], m0[ADAN_BATCH], d0[ADAN_BATCH], v0[ADAN_BATCH],
        p0[ADAN_BATCH];
    float* pp[ADAN_BATCH];
    int sg[ADAN_BATCH];
#pragma unroll
    for (int k = 0; k < ADAN_BATCH; ++k) {
      const int64_t i = i0 + k * stride;
      const int64_t ii = i < total ? i : 0;
      int seg = 0;
#pragma unroll
      for (int q = 1; q < 7; ++q) seg += ii >= A.off[q] ? 1 : 0;
      sg[k] = seg;
      pp[k] = A.param[seg] + (ii - A.off[seg]);
      g[k] = grads[ii];
      pg[k] = npg[ii];
      m0[k] = m[ii];
      d0[k] = dd[ii];
      v0[k] = v[ii];
      p0[k] = *pp[k];
    }
#pragma unroll
    for (int k = 0; k < ADAN_BATCH; ++k) {
      const int64_t i = i0 + k * stride;
      if (i >= total) continue;
      const int seg = sg[k];
      bool first = step == 1;
      if (seg == A.curv_seg) {
        if (!A.curv_enabled || itr < A.curv_warmup) continue;  // no gradient: untouched
        first = first || itr == A.curv_warmup;                 // state created at unfreeze
      }
      const float lr = seg == 0 ? lr_means : A.lr[seg];
      const float gi = g[k] * s_coef;
      const float prev = first ? -gi : pg[k];
      const float diff = prev + gi;
      const float mi = m0[k] * A.b1 + (1.0f - A.b1) * gi;
      const float di = d0[k] * A.b2 + (1.0f - A.b2) * diff;
      const float u = diff * A.b2 + gi;
      const float vi = v0[k] * A.b3 + (1.0f - A.b3) * u * u;
      const float denom = sqrtf(vi) / sbc3 + A.eps;
      float pv = p0[k];
      pv -= (lr / bc1) * mi / denom;
      pv -= (lr * A.b2 / bc2) * di / denom;
      *pp[k] = pv;
      m[i] = mi;
      v[i] = vi;
      dd[i] = di;
      npg[i] = -gi;
    }
  }
}

__global__ void k_advance(int32_t* iteration, const uint32_t* skip) {
  if (!(skip && skip[0])) iteration[0] += 1;
}

}  // namespace cvqpg

using namespace cvqpg;

extern "C" int cvqpg_adan_step(float* const* params, const int64_t* sizes, const float* grads,
                               float* state, const float* sumsq_partials, int n_partials,
                               const float* group_lr, float lr0, float eta_min, int num_itrs,
                               float beta1, float beta2, float beta3, float eps, float max_norm,
                               int curv_warmup, int curv_enabled, int32_t* iteration,
                               const uint32_t* skip_flag, void* stream) {
  CVQPG_CHECK_ARG(params && sizes && grads && state && sumsq_partials && group_lr && iteration,
                  "NULL argument pointer");
  AdanParams A;
  A.off[0] = 0;
  for (int k = 0; k < 7; ++k) {
    CVQPG_CHECK_ARG(params[k] != nullptr && sizes[k] >= 0, "parameter %d invalid", k);
    A.param[k] = params[k];
    A.off[k + 1] = A.off[k] + sizes[k];
    A.lr[k] = group_lr[k];
  }
  A.lr0 = lr0;
  A.eta_min = eta_min;
  A.num_itrs = num_itrs;
  A.b1 = beta1;
  A.b2 = beta2;
  A.b3 = beta3;
  A.eps = eps;
  A.max_norm = max_norm;
  A.curv_seg = 6;
  A.curv_warmup = curv_warmup;
  A.curv_enabled = curv_enabled;
  hipStream_t s = (hipStream_t)stream;
  const int64_t total = A.off[7];
  if (total > 0)
    hipLaunchKernelGGL(k_adan, dim3(std::min<int64_t>(ceil_div(total, 256), 1024)), dim3(256), 0,
                       s, A, grads, state, sumsq_partials, n_partials, iteration, skip_flag);
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, s, iteration, skip_flag);
  CVQPG_LAUNCH_CHECK();
  return 0;
}
