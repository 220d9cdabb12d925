// Adjoint of one [Conv] -> ReLU -> [AvgPool] stage, per pre-pool element.
//
// Forward of the element (h, h', w, w') of pair p (stax/linear.py:1387-1398,
// stax/elementwise.py:444-455, stax/linear.py:3551-3571):
//   c  = w_var * aN + b_var              (aN = stencil mean of nngp input)
//   t  = w_var * aT + c   (has_t)  or  t = c   (first stage: ntk input 0)
//   s  = sqrt(max(qa qb - c^2, 0)), theta = atan2(s, c), ds = (pi-theta)/2pi
//   J  = (s + (pi - theta) c) / 2pi,     TD = t ds
//   nngp_out = pool(J), ntk_out = pool(TD)   (pool = mean over p x p x p x p)
// For ABRelu(a, b) the map is J' = k1 J + k0 c, ds' = k1 ds + k0
// (k1 = (a - b)^2, k0 = a b; elementwise.py:444-455), and J, ds below stand
// for J', ds'.
// Given the pooled cotangents gN, gT this writes, per element,
//   gaN = w_var (gc + gt)   gaT = w_var gt        (cotangents of aN, aT)
//   gqa, gqb                                     (cotangents of qa, qb)
//   gwv = (gc + gt) aN + gt aT,  gbv = gc + gt     (of w_var, b_var; summed
//                                                  per workgroup into gwb)
// with gt = gTD ds and gc = gJ ds + gTD t d(ds)/dc.  The derivative rules at
// s = 0 follow the reference's custom JVPs (_sqrt_jvp and _arctan2_jvp,
// stax/elementwise.py:1283-1310): the tangent of s is zero there, so the
// terms in 1/s vanish.  Pairs of identical images (a self tile, n2 == 0, or
// the pairs (a, b) of a cross tile with a + self_off == b) have rho = 1
// exactly on the spatial diagonal (h = h', w = w'); there s = 0 is used
// instead of its rounded value, which gives the exact total derivative of
// K(x, x).  The stencil mean is self-adjoint, so the cotangents of the stage
// inputs are the forward stencil applied to gaN, gaT.
#include "nta_common.h"

namespace nta {

__global__ void k_relu_bwd(const float* __restrict__ aN,
                           const float* __restrict__ aT, int n1, int n2,
                           int S, int pool, int self_off, float w_var,
                           float b_var, float k1, float k0,
                           const float* __restrict__ q1,
                           const float* __restrict__ q2,
                           const float* __restrict__ gN,
                           const float* __restrict__ gT,
                           float* __restrict__ gaN, float* __restrict__ gaT,
                           float* __restrict__ gqa, float* __restrict__ gqb,
                           double* __restrict__ gwb) {
  __shared__ double red[2][4];
  const bool self = n2 == 0;
  double sw = 0.0, sb = 0.0;
  const uint64_t T = self ? (uint64_t)n1 : (uint64_t)n1 * n2;
  const uint64_t per = (uint64_t)S * S * S * S;
  const int So = S / pool;
  const uint64_t per_o = (uint64_t)So * So * So * So;
  const float inv_pool = 1.0f / (float)(pool * pool * pool * pool);
  const uint64_t total = T * per;
  for (uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
       idx < total; idx += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = idx % per;
    const uint64_t p = idx / per;
    const int w2 = (int)(r % S);
    const int w1 = (int)((r / S) % S);
    const int h2 = (int)((r / ((uint64_t)S * S)) % S);
    const int h1 = (int)(r / ((uint64_t)S * S * S));
    const uint64_t i = self ? p : p / n2;
    const uint64_t j = self ? p : p % n2;
    const float qa = q1[(i * S + h1) * S + w1];
    const float qb = (self ? q1 : q2)[(j * S + h2) * S + w2];
    const float an = aN[idx];
    const float at = aT ? aT[idx] : 0.0f;
    const float c = fmaf(w_var, an, b_var);
    const float t = aT ? fmaf(w_var, at, c) : c;
    const bool same = (self || (long long)i + self_off == (long long)j) &&
                      h1 == h2 && w1 == w2;
    const float d = prod_minus_sq(qa, qb, c);
    const float s = same ? 0.0f : sqrtf(fmaxf(d, 0.0f));
    const float theta = (s == 0.0f && c == 0.0f) ? 0.5f * kPi : atan2f(s, c);
    const float ds = fmaf(k1, (kPi - theta) * kInv2Pi, k0);   // dot_sigma'
    const uint64_t o = p * per_o +
                       (((uint64_t)(h1 / pool) * So + h2 / pool) * So +
                        w1 / pool) * So + w2 / pool;
    const float gj = gN ? gN[o] * inv_pool : 0.0f;
    const float gtd = gT ? gT[o] * inv_pool : 0.0f;
    // J' = k1 J + k0 c derivatives: dJ'/dc = ds', dJ'/dq = k1 s / (4 pi q)
    float dj_qa = qa > 0.0f ? k1 * s * (0.25f / kPi) / qa : 0.0f;
    float dj_qb = qb > 0.0f ? k1 * s * (0.25f / kPi) / qb : 0.0f;
    // ds' derivatives (zero where s = 0): d(ds')/dc = k1 / (2 pi s),
    // d(ds')/dq = -k1 c / (4 pi s q)
    float dd_c = 0.0f, dd_qa = 0.0f, dd_qb = 0.0f;
    if (s > 0.0f) {
      dd_c = k1 * kInv2Pi / s;
      dd_qa = qa > 0.0f ? -k1 * c * (0.25f / kPi) / (s * qa) : 0.0f;
      dd_qb = qb > 0.0f ? -k1 * c * (0.25f / kPi) / (s * qb) : 0.0f;
    }
    const float gt = gtd * ds;
    const float gc = gj * ds + gtd * t * dd_c;
    const float gct = gc + gt;
    gaN[idx] = w_var * gct;
    if (gaT) gaT[idx] = aT ? w_var * gt : 0.0f;
    gqa[idx] = gj * dj_qa + gtd * t * dd_qa;
    gqb[idx] = gj * dj_qb + gtd * t * dd_qb;
    sw += (double)(gct * an + gt * at);
    sb += (double)gct;
  }
  // Per-workgroup partial sums of the d/dw_var and d/db_var terms (fixed
  // order: wave shuffles, then the four waves).
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    sw += __shfl_xor(sw, o);
    sb += __shfl_xor(sb, o);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wave] = sw;
    red[1][wave] = sb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    gwb[2 * blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    gwb[2 * blockIdx.x + 1] =
        (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// Workgroups of a k_relu_bwd launch (= partial-sum slots of gwb).
static long long relu_bwd_blocks(int n1, int n2, int S) {
  const uint64_t T = n2 == 0 ? (uint64_t)n1 : (uint64_t)n1 * n2;
  const uint64_t total = T * S * S * S * S;
  uint64_t g = (total + 255) / 256;
  if (g > (1u << 16)) g = 1u << 16;
  return (long long)g;
}

}  // namespace nta

using namespace nta;

extern "C" long long nta_relu_bwd_parts(int n1, int n2, int S) {
  if (n1 < 0 || n2 < 0 || S <= 0) return -1;
  if (n1 == 0) return 0;
  return relu_bwd_blocks(n1, n2, S);
}

extern "C" int nta_relu_bwd(const float* aN, const float* aT, int n1, int n2,
                            int S, int pool, int self_off, float w_var,
                            float b_var, float act_k1, float act_k0,
                            const float* q1, const float* q2, const float* gN,
                            const float* gT, float* gaN, float* gaT,
                            float* gqa, float* gqb, double* gwb,
                            void* stream) {
  if (n1 < 0 || n2 < 0 || S <= 0 || (pool != 1 && pool != 2) ||
      S % pool != 0)
    return set_error(NTA_EINVAL, "nta_relu_bwd: bad n1=%d n2=%d S=%d "
                     "pool=%d", n1, n2, S, pool);
  if (n1 == 0) return NTA_OK;
  if (!aN || !q1 || (n2 > 0 && !q2) || !gaN || !gqa || !gqb || !gwb)
    return set_error(NTA_EINVAL, "nta_relu_bwd: null pointer");
  const long long g = relu_bwd_blocks(n1, n2, S);
  hipLaunchKernelGGL(k_relu_bwd, dim3((unsigned)g), dim3(256), 0,
                     (hipStream_t)stream, aN, aT, n1, n2, S, pool, self_off,
                     w_var, b_var, act_k1, act_k0, q1, q2, gN, gT, gaN, gaT,
                     gqa, gqb, gwb);
  return check_launch("nta_relu_bwd");
}
