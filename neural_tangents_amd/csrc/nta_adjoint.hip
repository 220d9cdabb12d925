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
// and, summed over the partner axes, the cotangents of the normalisers,
//   gqa[p][h][w] = sum over (h', w'),  gqb[p][h'][w'] = sum over (h, w),
// and the d/dw_var, d/db_var terms gwv = (gc + gt) aN + gt aT,
// gbv = gc + gt (summed per workgroup into gwb),
// with gt = gTD ds and gc = gJ ds + gTD t d(ds)/dc.  The derivative rules at
// s = 0 follow the reference's custom JVPs (_sqrt_jvp and _arctan2_jvp,
// stax/elementwise.py:1283-1310): the tangent of s is zero there, so the
// terms in 1/s vanish.  Pairs of identical images (a self tile, n2 == 0, or
// the pairs (a, b) of a cross tile with a + self_off == b) have rho = 1
// exactly on the spatial diagonal (h = h', w = w'); there s = 0 is used
// instead of its rounded value, which gives the exact total derivative of
// K(x, x).  The stencil mean is self-adjoint, so the cotangents of the stage
// inputs are the forward stencil applied to gaN, gaT.
//
// Sides 8, 16 and 32 (the Myrtle stages) run k_relu_bwd_rows: one workgroup
// per (pair, h), whose threads keep their (w, w') positions and sweep h', so
// the index arithmetic is per thread, not per element, and both normaliser
// sums are formed on chip (gqa completely, gqb up to a sum over h of
// S^3 floats per pair).  Other sides run the per-element k_relu_bwd, whose
// per-element normaliser terms are summed by k_bwd_qsum.
#include "nta_common.h"

namespace nta {

// atan2(y, x) for y >= 0, in [0, pi]: the smaller of |x|, y over the
// larger, t in [0, 1], is reduced to |u| <= tan(pi / 8) by
// u = (t - 1) / (t + 1) above tan(pi / 8), and atan(u) = u + u z P(z)
// (z = u^2, P of degree 3) -- the single-precision arctangent of the Cephes
// library (S. L. Moshier) -- before the octant is restored.  Absolute error
// below 3e-7 (hardware reciprocals).  (0, 0) gives 0.
__device__ __forceinline__ float atan2_pos(float y, float x) {
  const float ax = fabsf(x);
  const float mx = fmaxf(ax, y), mn = fminf(ax, y);
  const float t = mx > 0.0f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
  const bool big = t > 0.41421356237f;
  const float u = big ? (t - 1.0f) * __builtin_amdgcn_rcpf(t + 1.0f) : t;
  const float z = u * u;
  const float pz = fmaf(fmaf(fmaf(8.05374449538e-2f, z, -1.38776856032e-1f),
                             z, 1.99777106478e-1f),
                        z, -3.33329491539e-1f);
  float r = fmaf(pz * z, u, u);
  if (big) r += 0.5f * 0.5f * kPi;
  if (y > ax) r = 0.5f * kPi - r;
  if (x < 0.0f) r = kPi - r;
  return r;
}

// Cotangents of one element.  rqa, rqb: k1 / (4 pi q) (0 where q <= 0).
// Square root and reciprocals on the hardware instructions (~1 ulp): the
// cotangents feed gradients, not the kernels themselves.
struct BwdElem {
  float gct;     // gc + gt
  float gt;      // gTD ds
  float gq_a;    // of qa
  float gq_b;    // of qb
};

__device__ __forceinline__ BwdElem bwd_elem(float an, float at, bool has_t,
                                            float qa, float qb, float rqa,
                                            float rqb, bool same, float gj,
                                            float gtd, float w_var,
                                            float b_var, float k1, float k0) {
  const float c = fmaf(w_var, an, b_var);
  const float t = has_t ? fmaf(w_var, at, c) : c;
  const float d = prod_minus_sq(qa, qb, c);
  const float s = same ? 0.0f : fsqrt(fmaxf(d, 0.0f));
  const float th = atan2_pos(s, c);                    // no branch
  const float theta = (s == 0.0f && c == 0.0f) ? 0.5f * kPi : th;
  const float ds = fmaf(k1, (kPi - theta) * kInv2Pi, k0);   // dot_sigma'
  // d(ds')/dc = k1 / (2 pi s) and d(ds')/dq = -k1 c / (4 pi s q) vanish at
  // s = 0; dJ'/dc = ds', dJ'/dq = k1 s / (4 pi q)
  const float rs = s > 0.0f ? __builtin_amdgcn_rcpf(s) : 0.0f;
  const float gtt = gtd * t;
  const float gt = gtd * ds;
  const float gc = fmaf(gtt, k1 * kInv2Pi * rs, gj * ds);
  const float v = fmaf(gj, s, -gtt * (c * rs));
  BwdElem e;
  e.gct = gc + gt;
  e.gt = gt;
  e.gq_a = rqa * v;
  e.gq_b = rqb * v;
  return e;
}

__device__ __forceinline__ float bwd_rq(float q, float k1) {
  return q > 0.0f ? k1 * (0.25f / kPi) * __builtin_amdgcn_rcpf(q) : 0.0f;
}

// Sum of a workgroup's (sw, sb) in a fixed order (wave shuffles, then the
// four waves) into gwb[2 slot], gwb[2 slot + 1].  Every thread calls.
__device__ __forceinline__ void bwd_wb_sums(double sw, double sb,
                                            double* __restrict__ gwb,
                                            long long slot,
                                            double (&red)[2][4]) {
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
    gwb[2 * slot] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    gwb[2 * slot + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// Side S in {8, 16, 32}: workgroup (p, h) = (blockIdx.x, blockIdx.y) of 256
// threads.  Thread t holds w' = t % S and the w of its lane row r = t / S:
// w = r % kW1 + kW1 u (u < kU), and sweeps h' = r / kW1 + kH it.  Per h',
// the gqb terms of a lane row are summed over the thread's u and the wave's
// rows (shuffles) into lb[group][h'][w']; after the sweep the groups are
// added into gqb_part[p][h][h'][w'] and the gqa terms (summed over the
// sweep in registers, then over w' by shuffles) into gqa[p][h][w].
// gram_scale != 0 (cross tiles): gaN goes to the patch-Gram layout
// [n1][S][S][n2][S][S] = (a, h, w, b, h', w'), times gram_scale.
// HT, HN, HG: aT, gN, gT given (else zero).  With the presence of the
// inputs fixed at compile time the loop has no branches, so the loads of a
// thread's elements are all in flight together.
template <int S, bool HT, bool HN, bool HG>
__global__ __launch_bounds__(256) void k_relu_bwd_rows(
    const float* __restrict__ aN, const float* __restrict__ aT, int n2,
    int pool_sh, int self_off, float w_var, float b_var, float k1, float k0,
    const float* __restrict__ q1, const float* __restrict__ q2,
    const float* __restrict__ gN, const float* __restrict__ gT,
    float gram_scale, float* __restrict__ gaN, float* __restrict__ gaT,
    float* __restrict__ gqa, float* __restrict__ gqb_part,
    double* __restrict__ gwb) {
  static_assert(S == 8 || S == 16 || S == 32, "side");
  constexpr int kRows = 256 / S;                  // lane rows
  constexpr int kW1 = kRows < S ? kRows : S;      // w per u step
  constexpr int kU = S / kW1;                     // w per thread
  constexpr int kH = kRows / kW1;                 // h' planes at once
  constexpr int kG = S >= 16 ? 4 : 1;             // groups of the gqb sums
  __shared__ float lb[kG * S * S];
  __shared__ float la[kH * S];
  __shared__ double red[2][4];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int p = blockIdx.x, h1 = blockIdx.y;
  const bool self = n2 == 0;
  const int i = self ? p : p / n2;
  const int j = self ? p : p - i * n2;
  const bool self_pair = self || (long long)i + self_off == (long long)j;
  const int w2 = t & (S - 1);
  const int r = t / S;
  const int w1g = r % kW1;
  const int hoff = r / kW1;
  const int So = S >> pool_sh;
  const float inv_pool = pool_sh ? 1.0f / 16.0f : 1.0f;
  constexpr int S2 = S * S;
  // the workgroup's slabs (uniform bases; 32-bit offsets inside them)
  const long long slab = (long long)p * S2 * S2 + (long long)h1 * S2 * S;
  const float* aNs = aN + slab;
  const float* aTs = HT ? aT + slab : nullptr;
  float* gaTs = gaT ? gaT + slab : nullptr;
  const long long oslab = (long long)p * So * So * So * So +
                          (long long)(h1 >> pool_sh) * So * So * So;
  const float* gNs = HN ? gN + oslab : nullptr;
  const float* gTs = HG ? gT + oslab : nullptr;
  const bool gram = gram_scale != 0.0f;
  const float wg = gram ? w_var * gram_scale : w_var;
  // gaN: stage layout as the inputs, or (a, h, w, b, h', w')
  float* gaNs = gram ? gaN + ((long long)i * S + h1) * S * n2 * S2 +
                           (long long)j * S2
                     : gaN + slab;
  const int gw = gram ? n2 * S2 : S;      // gaN stride of w
  const int gh = gram ? S : S2;           // and of h'
  const float* qa_row = q1 + ((long long)i * S + h1) * S;
  const float* qb_img = (self ? q1 : q2) + (long long)j * S2;
  float qa[kU], rqa[kU], acc_a[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int w1 = w1g + kW1 * u;
    qa[u] = qa_row[w1];
    rqa[u] = bwd_rq(qa[u], k1);
    acc_a[u] = 0.0f;
  }
  double sw = 0.0, sb = 0.0;
  constexpr int kUnroll = kU >= 4 ? 1 : 2;        // S = 32: 4 elements
#pragma unroll kUnroll
  for (int it = 0; it < S / kH; ++it) {
    const int h2 = hoff + kH * it;
    const float qb = qb_img[h2 * S + w2];
    const float rqb = bwd_rq(qb, k1);
    const unsigned orow = (unsigned)((h2 >> pool_sh) * So * So +
                                     (w2 >> pool_sh));
    float an[kU], at[kU], gj[kU], gtd[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {              // all loads first
      const int w1 = w1g + kW1 * u;
      const unsigned e = (unsigned)((h2 * S + w1) * S + w2);
      const unsigned o = orow + (unsigned)((w1 >> pool_sh) * So);
      an[u] = aNs[e];
      at[u] = HT ? aTs[e] : 0.0f;
      gj[u] = HN ? gNs[o] * inv_pool : 0.0f;
      gtd[u] = HG ? gTs[o] * inv_pool : 0.0f;
    }
    float gbs = 0.0f;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int w1 = w1g + kW1 * u;
      const unsigned e = (unsigned)((h2 * S + w1) * S + w2);
      const bool same = self_pair && h1 == h2 && w1 == w2;
      const BwdElem el = bwd_elem(an[u], at[u], HT, qa[u], qb, rqa[u], rqb,
                                  same, gj[u], gtd[u], w_var, b_var, k1, k0);
      gaNs[(unsigned)(w1 * gw + h2 * gh + w2)] = wg * el.gct;
      if (HT) gaTs[e] = w_var * el.gt;
      else if (gaTs) gaTs[e] = 0.0f;
      acc_a[u] += el.gq_a;
      gbs += el.gq_b;
      sw += (double)(el.gct * an[u] + el.gt * at[u]);
      sb += (double)el.gct;
    }
    // over the other rows of the wave (w = w1g + ..., same h', w')
#pragma unroll
    for (int m = S; m < 64; m <<= 1) gbs += __shfl_xor(gbs, m);
    if (lane < S) lb[(kG == 4 ? wave : 0) * S * S + h2 * S + w2] = gbs;
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
#pragma unroll
    for (int m = 1; m < S; m <<= 1) acc_a[u] += __shfl_xor(acc_a[u], m);
  }
  if (w2 == 0) {
#pragma unroll
    for (int u = 0; u < kU; ++u) la[hoff * S + w1g + kW1 * u] = acc_a[u];
  }
  bwd_wb_sums(sw, sb, gwb, (long long)p * S + h1, red);  // has a barrier
  if (t < S) {
    float v = 0.0f;
#pragma unroll
    for (int h = 0; h < kH; ++h) v += la[h * S + t];
    gqa[((long long)p * S + h1) * S + t] = v;
  }
  for (int k = t; k < S * S; k += 256) {
    float v = 0.0f;
#pragma unroll
    for (int g = 0; g < kG; ++g) v += lb[g * S * S + k];
    gqb_part[((long long)p * S + h1) * S2 + k] = v;
  }
}

// gqb[p][k] = sum over h of part[p][h][k] (k < S^2), in order of h (the
// S loads of a thread unrolled, so they are in flight together).
template <int S>
__global__ void k_bwd_sum_h(const float* __restrict__ part, long long T,
                            float* __restrict__ gqb) {
  constexpr long long S2 = (long long)S * S;
  const long long n = T * S2;
  for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < n;
       x += (long long)gridDim.x * blockDim.x) {
    const long long p = x / S2, k = x - p * S2;
    const float* src = part + p * S2 * S + k;
    float t[S];
#pragma unroll
    for (int h = 0; h < S; ++h) t[h] = src[h * S2];
    float v = 0.0f;
#pragma unroll
    for (int h = 0; h < S; ++h) v += t[h];
    gqb[x] = v;
  }
}

// Any side: one element per thread of a grid-stride loop; the per-element
// gqa, gqb terms go to ea, eb ([T][S][S][S][S]) for k_bwd_qsum.
__global__ void k_relu_bwd(const float* __restrict__ aN,
                           const float* __restrict__ aT, int n1, int n2,
                           int S, int pool, int self_off, float w_var,
                           float b_var, float k1, float k0,
                           const float* __restrict__ q1,
                           const float* __restrict__ q2,
                           const float* __restrict__ gN,
                           const float* __restrict__ gT, float gram_scale,
                           float* __restrict__ gaN, float* __restrict__ gaT,
                           float* __restrict__ ea, float* __restrict__ eb,
                           double* __restrict__ gwb) {
  __shared__ double red[2][4];
  const bool self = n2 == 0;
  const bool has_t = aT != nullptr;
  const bool gram = gram_scale != 0.0f;
  const float wg = gram ? w_var * gram_scale : w_var;
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
    const float at = has_t ? aT[idx] : 0.0f;
    const bool same = (self || (long long)i + self_off == (long long)j) &&
                      h1 == h2 && w1 == w2;
    const uint64_t o = p * per_o +
                       (((uint64_t)(h1 / pool) * So + h2 / pool) * So +
                        w1 / pool) * So + w2 / pool;
    const float gj = gN ? gN[o] * inv_pool : 0.0f;
    const float gtd = gT ? gT[o] * inv_pool : 0.0f;
    const BwdElem el = bwd_elem(an, at, has_t, qa, qb, bwd_rq(qa, k1),
                                bwd_rq(qb, k1), same, gj, gtd, w_var, b_var,
                                k1, k0);
    const uint64_t go =
        gram ? (((i * S + h1) * S + w1) * n2 + j) * S * S + h2 * S + w2 : idx;
    gaN[go] = wg * el.gct;
    if (gaT) gaT[idx] = has_t ? w_var * el.gt : 0.0f;
    ea[idx] = el.gq_a;
    eb[idx] = el.gq_b;
    sw += (double)(el.gct * an + el.gt * at);
    sb += (double)el.gct;
  }
  bwd_wb_sums(sw, sb, gwb, blockIdx.x, red);
}

// gqa[p][x][y] = sum over (h', w') of ea[p][x][h'][y][w'] and
// gqb[p][x][y] = sum over (h, w) of eb[p][h][x][w][y], in a fixed order.
__global__ void k_bwd_qsum(const float* __restrict__ ea,
                           const float* __restrict__ eb, long long T, int S,
                           float* __restrict__ gqa, float* __restrict__ gqb) {
  const long long S2 = (long long)S * S;
  const long long n = T * S2;
  for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < n;
       o += (long long)gridDim.x * blockDim.x) {
    const long long p = o / S2;
    const int x = (int)((o / S) % S), y = (int)(o % S);
    const float* a = ea + p * S2 * S2 + (long long)x * S2 * S + y * S;
    const float* b = eb + p * S2 * S2 + (long long)x * S2 + y;
    float va = 0.0f, vb = 0.0f;
    for (int u = 0; u < S; ++u)
      for (int v = 0; v < S; ++v) {
        va += a[(long long)u * S2 + v];                 // (h' = u, w' = v)
        vb += b[(long long)u * S2 * S + (long long)v * S];  // (h = u, w = v)
      }
    gqa[o] = va;
    gqb[o] = vb;
  }
}

// acc[0], acc[1] += the sums of the (d/dw_var, d/db_var) partial slots
// gwb[2 s], gwb[2 s + 1], s < parts, in a fixed order (one workgroup; a
// thread's slots t, t + 256, ... loaded eight at a time, added in order).
__global__ __launch_bounds__(256) void k_bwd_acc_wb(
    const double* __restrict__ gwb, long long parts, double* __restrict__ acc) {
  __shared__ double red[2][256];
  const int t = threadIdx.x;
  double w = 0.0, b = 0.0;
  long long q = t;
  for (; q + 7 * 256 < parts; q += 8 * 256) {
    double vw[8], vb[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      vw[u] = gwb[2 * (q + u * 256)];
      vb[u] = gwb[2 * (q + u * 256) + 1];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      w += vw[u];
      b += vb[u];
    }
  }
  for (; q < parts; q += 256) {
    w += gwb[2 * q];
    b += gwb[2 * q + 1];
  }
  red[0][t] = w;
  red[1][t] = b;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      red[0][t] += red[0][t + h];
      red[1][t] += red[1][t + h];
    }
    __syncthreads();
  }
  if (t == 0) {
    acc[0] += red[0][0];
    acc[1] += red[1][0];
  }
}

// acc[o][k] += sum over b < nb of src[(o so + b sb) S2 + k] (k < S2), in
// order of b: a normaliser cotangent summed over the partner images.
__global__ __launch_bounds__(256) void k_bwd_acc_q(
    const float* __restrict__ src, int no, int nb, long long so,
    long long sb, int S2, float* __restrict__ acc) {
  const long long n = (long long)no * S2;
  for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < n;
       x += (long long)gridDim.x * blockDim.x) {
    const long long o = x / S2;
    const int k = (int)(x - o * S2);
    float v = 0.0f;
    for (int b = 0; b < nb; ++b) v += src[(o * so + b * sb) * S2 + k];
    acc[x] += v;
  }
}

// k_bwd_acc_q for many partners (a tile of one image against a row of
// others): a workgroup owns 64 consecutive k of one o, its four waves sum
// the partners b = w, w + 4, ... and the four sums are added in a fixed
// order.
__global__ __launch_bounds__(256) void k_bwd_acc_q4(
    const float* __restrict__ src, int no, int nb, long long so,
    long long sb, int S2, float* __restrict__ acc) {
  __shared__ float red[3][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long long kblocks = (S2 + 63) / 64, nblk = (long long)no * kblocks;
  for (long long blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const long long o = blk / kblocks;
    const int k = (int)(blk - o * kblocks) * 64 + lane;
    float v = 0.0f;
    if (k < S2) {
      const float* s = src + o * so * S2 + k;
      const long long step = sb * S2;
#pragma unroll 4
      for (int b = w; b < nb; b += 4) v += s[b * step];
    }
    if (w) red[w - 1][lane] = v;
    __syncthreads();
    if (w == 0 && k < S2)
      acc[o * S2 + k] += (v + red[0][lane]) + (red[1][lane] + red[2][lane]);
    __syncthreads();                  // red is rewritten by the next block
  }
}


static bool bwd_rows(int S) { return S == 8 || S == 16 || S == 32; }

static long long bwd_pairs(int n1, int n2) {
  return n2 == 0 ? (long long)n1 : (long long)n1 * n2;
}

// Workgroups of a k_relu_bwd launch (= partial-sum slots of gwb).
static long long relu_bwd_blocks(int n1, int n2, int S) {
  const uint64_t total = (uint64_t)bwd_pairs(n1, n2) * S * S * S * S;
  uint64_t g = (total + 255) / 256;
  if (g > (1u << 16)) g = 1u << 16;
  return (long long)g;
}

static unsigned grid_for(long long n) {
  long long g = (n + 255) / 256;
  return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

// acc[o] += sum over b of src[o so + b sb] (rows of S2 floats).
static void acc_q(const float* src, int no, int nb, long long so,
                  long long sb, int S2, float* acc, hipStream_t st) {
  if (nb >= 8) {
    const long long nblk = (long long)no * ((S2 + 63) / 64);
    hipLaunchKernelGGL(k_bwd_acc_q4,
                       dim3((unsigned)(nblk < 65536 ? nblk : 65536)),
                       dim3(256), 0, st, src, no, nb, so, sb, S2, acc);
  } else {
    hipLaunchKernelGGL(k_bwd_acc_q, dim3(grid_for((long long)no * S2)),
                       dim3(256), 0, st, src, no, nb, so, sb, S2, acc);
  }
}

}  // namespace nta

using namespace nta;

extern "C" long long nta_relu_bwd_parts(int n1, int n2, int S) {
  if (n1 < 0 || n2 < 0 || S <= 0) return -1;
  if (n1 == 0) return 0;
  return bwd_rows(S) ? bwd_pairs(n1, n2) * S : relu_bwd_blocks(n1, n2, S);
}

extern "C" long long nta_relu_bwd_work_bytes(int n1, int n2, int S) {
  if (n1 < 0 || n2 < 0 || S <= 0) return -1;
  const long long s3 = (long long)S * S * S;
  return bwd_rows(S) ? bwd_pairs(n1, n2) * s3 * 4
                     : 2 * bwd_pairs(n1, n2) * s3 * S * 4;
}

extern "C" int nta_relu_bwd(const float* aN, const float* aT, int n1, int n2,
                            int S, int pool, int self_off, float w_var,
                            float b_var, float act_k1, float act_k0,
                            const float* q1, const float* q2, const float* gN,
                            const float* gT, float gram_scale, float* gaN,
                            float* gaT, float* gqa, float* gqb, double* gwb,
                            void* work, long long work_bytes, void* stream) {
  if (n1 < 0 || n2 < 0 || S <= 0 || (pool != 1 && pool != 2) ||
      S % pool != 0)
    return set_error(NTA_EINVAL, "nta_relu_bwd: bad n1=%d n2=%d S=%d "
                     "pool=%d", n1, n2, S, pool);
  if (gram_scale != 0.0f && n2 == 0)
    return set_error(NTA_EINVAL, "nta_relu_bwd: the patch-Gram layout "
                     "(gram_scale != 0) needs a cross tile (n2 > 0)");
  if (n1 == 0) return NTA_OK;
  if (!aN || !q1 || (n2 > 0 && !q2) || !gaN || !gqa || !gqb || !gwb || !work)
    return set_error(NTA_EINVAL, "nta_relu_bwd: null pointer");
  const long long need = nta_relu_bwd_work_bytes(n1, n2, S);
  if (work_bytes < need)
    return set_error(NTA_EINVAL, "nta_relu_bwd: workspace of %lld bytes, "
                     "%lld needed", work_bytes, need);
  const long long T = bwd_pairs(n1, n2);
  if (T >= (1LL << 24))                 // grid.x * 256 work-items < 2^32
    return set_error(NTA_EINVAL, "nta_relu_bwd: %lld pairs in one tile", T);
  if (gram_scale != 0.0f && (long long)n2 * S * S * S >= (1LL << 31))
    return set_error(NTA_EINVAL, "nta_relu_bwd: patch-Gram layout with "
                     "n2 * S^3 >= 2^31");
  hipStream_t st = (hipStream_t)stream;
  float* part = (float*)work;
  if (bwd_rows(S)) {
    const int ps = pool == 2 ? 1 : 0;
    const dim3 grid((unsigned)T, (unsigned)S);
#define NTA_BWD_ROWS(SS, HT, HN, HG)                                        \
    hipLaunchKernelGGL((k_relu_bwd_rows<SS, HT, HN, HG>), grid, dim3(256),  \
                       0, st, aN, aT, n2, ps, self_off, w_var, b_var,       \
                       act_k1, act_k0, q1, q2, gN, gT, gram_scale, gaN,     \
                       gaT, gqa, part, gwb)
#define NTA_BWD_ROWS_G(SS, HT)                                              \
    if (gN && gT) NTA_BWD_ROWS(SS, HT, true, true);                         \
    else if (gN) NTA_BWD_ROWS(SS, HT, true, false);                         \
    else if (gT) NTA_BWD_ROWS(SS, HT, false, true);                         \
    else NTA_BWD_ROWS(SS, HT, false, false)
#define NTA_BWD_ROWS_S(SS)                                                  \
    if (aT) {                                                               \
      NTA_BWD_ROWS_G(SS, true);                                             \
    } else {                                                                \
      NTA_BWD_ROWS_G(SS, false);                                            \
    }
    if (S == 32) {
      NTA_BWD_ROWS_S(32);
    } else if (S == 16) {
      NTA_BWD_ROWS_S(16);
    } else {
      NTA_BWD_ROWS_S(8);
    }
#undef NTA_BWD_ROWS_S
#undef NTA_BWD_ROWS_G
#undef NTA_BWD_ROWS
    const dim3 sgrid(grid_for(T * S * S));
    if (S == 32)
      hipLaunchKernelGGL(k_bwd_sum_h<32>, sgrid, dim3(256), 0, st, part, T,
                         gqb);
    else if (S == 16)
      hipLaunchKernelGGL(k_bwd_sum_h<16>, sgrid, dim3(256), 0, st, part, T,
                         gqb);
    else
      hipLaunchKernelGGL(k_bwd_sum_h<8>, sgrid, dim3(256), 0, st, part, T,
                         gqb);
  } else {
    const long long per = (long long)S * S * S * S;
    const long long g = relu_bwd_blocks(n1, n2, S);
    hipLaunchKernelGGL(k_relu_bwd, dim3((unsigned)g), dim3(256), 0, st, aN,
                       aT, n1, n2, S, pool, self_off, w_var, b_var, act_k1,
                       act_k0, q1, q2, gN, gT, gram_scale, gaN, gaT, part,
                       part + T * per, gwb);
    hipLaunchKernelGGL(k_bwd_qsum, dim3(grid_for(T * S * S)), dim3(256), 0,
                       st, part, part + T * per, T, S, gqa, gqb);
  }
  return check_launch("nta_relu_bwd");
}

extern "C" int nta_relu_bwd_accumulate(const float* gqa, const float* gqb,
                                       const double* gwb, int n1, int n2,
                                       int S, double* wb_acc, float* gq1_acc,
                                       float* gq2_acc, void* stream) {
  if (n1 < 0 || n2 < 0 || S <= 0)
    return set_error(NTA_EINVAL, "nta_relu_bwd_accumulate: bad n1=%d n2=%d "
                     "S=%d", n1, n2, S);
  if (n1 == 0) return NTA_OK;
  if ((wb_acc && !gwb) || ((gq1_acc || gq2_acc) && (!gqa || !gqb)) ||
      (gq2_acc && n2 == 0))
    return set_error(NTA_EINVAL, "nta_relu_bwd_accumulate: null pointer or "
                     "a second-image accumulator for a self tile");
  hipStream_t st = (hipStream_t)stream;
  const int S2 = S * S;
  if (wb_acc)
    hipLaunchKernelGGL(k_bwd_acc_wb, dim3(1), dim3(256), 0, st, gwb,
                       nta_relu_bwd_parts(n1, n2, S), wb_acc);
  if (gq1_acc) {
    if (n2 == 0) {                     // self tile: both normalisers are q1
      acc_q(gqa, n1, 1, 1LL, 0LL, S2, gq1_acc, st);
      acc_q(gqb, n1, 1, 1LL, 0LL, S2, gq1_acc, st);
    } else {
      acc_q(gqa, n1, n2, (long long)n2, 1LL, S2, gq1_acc, st);
    }
  }
  // after gq1_acc (the two may alias for a tile of K(x, x))
  if (gq2_acc) acc_q(gqb, n2, n1, 1LL, (long long)n2, S2, gq2_acc, st);
  return check_launch("nta_relu_bwd_accumulate");
}
