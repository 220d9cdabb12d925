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

// Shape-specialised launchers (nta_fused.hip).  They return kNotHandled when
// the shape is outside their specialisation; the caller then launches the
// generic kernel.
constexpr int kNotHandled = 1 << 30;
int launch_stage_fast(const float* in_nngp, const float* in_ntk, int n1,
                      int n2, int S, int pool, int flags, float w_var,
                      float b_var, float k1, float k0, const float* q1,
                      const float* q2, float* out_nngp, float* out_ntk,
                      hipStream_t stream);
int launch_stencil_fast(const float* in, long long T, int S, float w_var,
                        float b_var, float* out, hipStream_t st);
int launch_first_fast(const float* x1, const float* x2, int n1, int n2, int S,
                      int C, int flags, float w_var, float b_var, float k1,
                      float k0, const float* q1, const float* q2,
                      float* nngp, float* ntk, hipStream_t stream);

// One-instruction square root on the device (v_sqrt_f32, ~1 ulp); the host
// pass uses the C library so the same code can be unit-tested on the CPU.
__host__ __device__ __forceinline__ float fsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}

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

// ABRelu(a, b) = a min(x, 0) + b max(x, 0) on kernels (elementwise.py:
// 423-477) is (a - b)^2 Relu + a b Id: nngp' = k1 J + k0 c and
// dot_sigma' = k1 dot_sigma + k0 with k1 = (a - b)^2, k0 = a b, where J and
// dot_sigma are the ReLU map's.  Relu is k1 = 1, k0 = 0 (bitwise the ReLU
// map).  LeakyRelu(alpha) = ABRelu(alpha, 1), Abs = ABRelu(-1, 1).
__device__ __forceinline__ float abrelu_map(float c, float q1, float q2,
                                            float k1, float k0,
                                            float* dot_sigma) {
  float d;
  const float j = relu_map(c, q1, q2, &d);
  *dot_sigma = fmaf(k1, d, k0);
  return fmaf(k1, j, k0 * c);
}

// Erf(a, b, c) on kernels (elementwise.py:49-108; the input scale b is
// applied to (c, q1, q2) by the caller): with d = (1 + 2 q1)(1 + 2 q2) - 4 c^2
// = 1 + 2 (q1 + q2) + 4 (q1 q2 - c^2) > 0,
//   nngp' = k1 (2/pi) atan2(2 c, sqrt(d)) + k0,
//   dot_sigma = k1 (4/pi) / sqrt(d)
// with k1 = a^2, k0 = c_shift^2 (a * Kernel + c, kernel.py:426-439).
__device__ __forceinline__ float erf_map(float c, float q1, float q2,
                                         float k1, float k0,
                                         float* dot_sigma) {
  const float d = fmaf(4.0f, prod_minus_sq(q1, q2, c),
                       fmaf(2.0f, q1 + q2, 1.0f));
  const float s = sqrtf(fmaxf(d, 0.0f));
  *dot_sigma = k1 * (2.0f / kPi) * 2.0f / s;
  return fmaf(k1 * (2.0f / kPi), atan2f(2.0f * c, s), k0);
}

// Gelu on kernels (elementwise.py:197-277): with p = q1 q2,
// P = (q1 + 1)(q2 + 1), D^2 = P - c^2 = (q1 q2 - c^2) + q1 + q2 + 1 > 0,
// t = atan2(c, D):
//   nngp' = ((c^2 + p D^2) / (P D) + c t) / (2 pi) + c / 4,
//   dot_sigma = c / (2 pi D) (1 / D^2 + (1 - p) / P + 1) + 1/4 + t / (2 pi).
__device__ __forceinline__ float gelu_map(float c, float q1, float q2,
                                          float* dot_sigma) {
  const float p = q1 * q2;
  const float P = (q1 + 1.0f) * (q2 + 1.0f);
  const float D2 = fmaxf(prod_minus_sq(q1, q2, c) + (q1 + q2 + 1.0f), 0.0f);
  const float D = sqrtf(D2);
  const float t = atan2f(c, D);
  *dot_sigma = c / (2.0f * kPi * D) * (1.0f / D2 + (1.0f - p) / P + 1.0f) +
               0.25f + t * kInv2Pi;
  return (fmaf(c, c, p * D2) / (P * D) + c * t) * kInv2Pi + 0.25f * c;
}

// The activation of a stage: flags & NTA_ACT_ERF / NTA_ACT_GELU select Erf or
// Gelu, otherwise ABRelu(k1, k0).
__device__ __forceinline__ float act_map(int flags, float c, float q1,
                                         float q2, float k1, float k0,
                                         float* dot_sigma) {
  if (flags & NTA_ACT_ERF) return erf_map(c, q1, q2, k1, k0, dot_sigma);
  if (flags & NTA_ACT_GELU) return gelu_map(c, q1, q2, dot_sigma);
  return abrelu_map(c, q1, q2, k1, k0, dot_sigma);
}

// Normalised form of the same map, used by the fused kernels.  With
// sq = sqrt(q) and rq = 1/sqrt(q) (0 when q == 0) per position,
//   rho = c rq1 rq2, theta = arccos(rho) (rho clamped to [-1, 1]),
//   nngp' = sq1 sq2 (sqrt(1 - rho^2) + (pi - theta) rho) / (2 pi),
//   dot_sigma = (pi - theta) / (2 pi).
// arccos(a), a in [0, 1], is Abramowitz & Stegun 4.4.46:
// sqrt(1 - a) * (a0 + a1 a + ... + a7 a^7), |error| <= 2e-8 (u = 1 - a).
// q == 0 gives rho = 0, theta = pi/2 and nngp' = 0, the reference's
// fill_zero convention (elementwise.py:446).
__host__ __device__ __forceinline__ float acos_pos(float a, float u) {
  float r = fsqrt(u);
  float p = -0.0012624911f;
  p = fmaf(p, a, 0.0066700901f);
  p = fmaf(p, a, -0.0170881256f);
  p = fmaf(p, a, 0.0308918810f);
  p = fmaf(p, a, -0.0501743046f);
  p = fmaf(p, a, 0.0889789874f);
  p = fmaf(p, a, -0.2145988016f);
  p = fmaf(p, a, 1.5707963050f);
  return r * p;
}

// The same form with a degree-5 polynomial, fitted here to the weighted
// minimax error of theta (Lawson's iteratively reweighted least squares on
// 2e5 points of [0, 1], weight sqrt(1 - a)): |error| <= 6.4e-7 in exact
// arithmetic, <= 8.4e-7 evaluated in FP32.  Two FMAs fewer than acos_pos;
// used by the fused Myrtle kernels, whose cost is the vector ALU.
__host__ __device__ __forceinline__ float acos_pos5(float a, float u) {
  float r = fsqrt(u);
  float p = -4.9111684784e-03f;
  p = fmaf(p, a, 2.0620046183e-02f);
  p = fmaf(p, a, -4.5927215368e-02f);
  p = fmaf(p, a, 8.8171049953e-02f);
  p = fmaf(p, a, -2.1454282105e-01f);
  p = fmaf(p, a, 1.5707956553e+00f);
  return r * p;
}

__host__ __device__ __forceinline__ float relu_map_norm(float c, float sqp,
                                                        float rqp,
                                                        float* dot_sigma) {
  float rho = c * rqp;
  float a = fminf(fabsf(rho), 1.0f);
  float u = 1.0f - a;
  float th = acos_pos(a, u);                  // arccos(|rho|)
  float pm = rho < 0.0f ? th : kPi - th;      // pi - arccos(rho)
  float sn = fsqrt(u * (2.0f - u));           // sqrt(1 - rho^2)
  *dot_sigma = pm * kInv2Pi;
  return fmaf(pm, rho, sn) * (sqp * kInv2Pi);
}

}  // namespace nta
