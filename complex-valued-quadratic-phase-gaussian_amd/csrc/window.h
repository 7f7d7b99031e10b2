// Guard-band window and amplitude/phase glue shared by the propagation and export
// kernels.
//
// Reference behaviour:
//   soft_window_pad   model_2d_gaussian_rgb.py:186-241 (Tukey ramps over the guard band)
//   reconstruct glue  train_2d_gaussian_rgb.py:89-95
#pragma once
#include <hip/hip_runtime.h>

namespace cvqpg {

// ---------------------------------------------------------------------------------
// glue: u = clamp(|p|, 0, 1) e^{i angle(p)}  (== p for |p| <= 1, p/|p| otherwise)
// ---------------------------------------------------------------------------------
__device__ __forceinline__ float2 glue(float2 p) {
  const float a2 = p.x * p.x + p.y * p.y;
  if (a2 > 1.0f) {
    const float inv = rsqrtf(a2);
    return make_float2(p.x * inv, p.y * inv);
  }
  return p;
}

// adjoint of glue for gradient g = dL/dRe u + i dL/dIm u (torch's convention, with the
// gradient of abs and angle defined as 0 at the origin)
__device__ __forceinline__ float2 glue_adjoint(float2 h, float2 g) {
  const float a2 = h.x * h.x + h.y * h.y;
  if (a2 == 0.0f) return make_float2(0.f, 0.f);
  if (a2 <= 1.0f) return g;
  const float a = sqrtf(a2);
  const float hx = h.x / a, hy = h.y / a;
  const float q = (g.y * hx - g.x * hy) / a;  // Im(g conj(hh)) / |h|
  return make_float2(-q * hy, q * hx);         // q * i * hh
}

// Tukey ramps 0.5 (1 - cos(pi k / n)) and 0.5 (1 + cos(pi k / n)) in their half-angle forms
// sin^2(pi k / 2n) and cos^2(pi k / 2n), which do not cancel near the start of a ramp, on the
// sine units (revolutions: k / 4n in [0, 1/4], no range reduction)
__device__ __forceinline__ float taper(int t, int total, int inner_start, int inner_size) {
  if (t < inner_start) {
    const float s = __builtin_amdgcn_sinf(0.25f * (float)(t + 1) / (float)inner_start);
    return s * s;
  }
  const int rs = inner_start + inner_size;
  if (t >= rs) {
    const int n_right = total - rs;
    const float c = __builtin_amdgcn_cosf(0.25f * (float)(t - rs + 1) / (float)n_right);
    return c * c;
  }
  return 1.0f;
}


}  // namespace cvqpg
