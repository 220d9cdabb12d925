// Shared device helpers and host-side status plumbing for libnta.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nta.h"

namespace nta {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInv2Pi = 0.15915494309189533577f;

// Sets the thread-local message returned by nta_last_error(); returns code.
int set_error(int code, const char* fmt, ...);
int check_launch(const char* what);

// q1*q2 - c*c with one rounding (Kahan's product-difference), so that pairs
// with q1*q2 == c*c in exact arithmetic give exactly 0.
__device__ __forceinline__ float prod_minus_sq(float q1, float q2, float c) {
  float w = c * c;
  float e = fmaf(-c, c, w);        // w - c*c, exact
  float f = fmaf(q1, q2, -w);      // q1*q2 - w, rounded once
  return f + e;
}

// ReLU arc-cosine map (ABRelu(0, 1), elementwise.py:444-455):
//   s = sqrt(max(q1 q2 - c^2, 0)); theta = atan2(s, c) (pi/2 when s = c = 0)
//   nngp' = (s + (pi - theta) c) / (2 pi); dot_sigma = (pi - theta) / (2 pi)
__device__ __forceinline__ float relu_map(float c, float q1, float q2,
                                          float* dot_sigma) {
  float d = prod_minus_sq(q1, q2, c);
  float s = sqrtf(fmaxf(d, 0.0f));
  float theta = (s == 0.0f && c == 0.0f) ? 0.5f * kPi : atan2f(s, c);
  float pm = kPi - theta;
  *dot_sigma = pm * kInv2Pi;
  return fmaf(pm, c, s) * kInv2Pi;
}

}  // namespace nta
