// Shape-specialised kernels for the Myrtle path (3x3 SAME Conv + ReLU
// [+ 2x2 AvgPool], square images of side 8/16/32).
//
// k_stage_fast: one fused Conv -> ReLU [-> AvgPool] stage on a tile of pairs.
//   A workgroup owns one diagonal (h - h' = const, in output planes) of JPW
//   pairs and sweeps it.  The 3x3 stencil is separable into a 3-tap sum along
//   the (w, w') diagonal of each (h, h') plane (done once per loaded plane
//   through an LDS exchange) and a 3-tap sum along the (h, h') diagonal (a
//   register sliding window over consecutive planes of the sweep), so every
//   input plane is read from HBM once per diagonal that needs it.  Each thread
//   owns a 2x2 (w, w') block, which makes the 2x2x2x2 pool register-local.
//
// k_first_fast: input covariance + first Conv + ReLU.  The stencil of the
//   input Gram is the Gram of zero-padded 3x3xC patches (K = 9 * 4 with C
//   padded to 4), computed per (h, h') plane with fp32 MFMA
//   (32x32x2 for S = 32, 16x16x4 for S = 16); the ReLU is the epilogue.
//
// Both remap blockIdx so that consecutive logical blocks (the diagonals or
// rows of one pair) land on the same XCD: dispatch is round-robin over the 8
// XCDs, each with its own L2.
#include "nta_common.h"

namespace nta {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float2 qnorm(float q) {
  float sq = fsqrt(fmaxf(q, 0.0f));
  float rq = q > 0.0f ? 1.0f / sq : 0.0f;
  return make_float2(sq, rq);
}

__device__ __forceinline__ int xcd_remap(int b, int nblk) {
  return (b & 7) * (nblk >> 3) + (b >> 3);
}

// (w, w')-diagonal 3-tap sum of K planes.  Plane k is (hh[k], hp[k]) of the
// pair state at base[k]; planes outside [0, S)^2 are zero.  Each thread owns
// the 2x2 block (2W.., 2Wp..).  ws[k][e], e = 2c + d, is element
// (2W + c, 2Wp + d).  Two barriers; every thread of the block must call.
template <int S, int K>
__device__ __forceinline__ void wpass(const float* const* base,
                                      const int* hh, const int* hp,
                                      float* __restrict__ lds, int W, int Wp,
                                      bool valid, float (&ws)[K][4]) {
  constexpr int LS = S + 2;
  constexpr int PL = LS * LS;
  bool ok[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    ok[k] = (unsigned)hh[k] < (unsigned)S && (unsigned)hp[k] < (unsigned)S;
    ws[k][0] = ws[k][1] = ws[k][2] = ws[k][3] = 0.0f;
    if (ok[k]) {
      float2 r0 = make_float2(0.0f, 0.0f), r1 = r0;
      if (valid) {
        const float* pl = base[k] + ((long long)hh[k] * S + hp[k]) * (S * S);
        r0 = *reinterpret_cast<const float2*>(pl + (2 * W) * S + 2 * Wp);
        r1 = *reinterpret_cast<const float2*>(pl + (2 * W + 1) * S + 2 * Wp);
      }
      float* d = lds + k * PL + (2 * W + 1) * LS + 2 * Wp + 1;
      d[0] = r0.x;
      d[1] = r0.y;
      d[LS] = r1.x;
      d[LS + 1] = r1.y;
      ws[k][0] = r0.x;
      ws[k][1] = r0.y;
      ws[k][2] = r1.x;
      ws[k][3] = r1.y;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (ok[k]) {
      const float* s0 = lds + k * PL + (2 * W + 1) * LS + 2 * Wp + 1;
      float v00 = ws[k][0], v01 = ws[k][1], v10 = ws[k][2], v11 = ws[k][3];
      float a_mm = s0[-LS - 1];       // (2W-1, 2Wp-1)
      float a_m0 = s0[-LS];           // (2W-1, 2Wp)
      float a_0m = s0[-1];            // (2W,   2Wp-1)
      float a_12 = s0[LS + 2];        // (2W+1, 2Wp+2)
      float a_21 = s0[2 * LS + 1];    // (2W+2, 2Wp+1)
      float a_22 = s0[2 * LS + 2];    // (2W+2, 2Wp+2)
      ws[k][0] = a_mm + v00 + v11;
      ws[k][1] = a_m0 + v01 + a_12;
      ws[k][2] = a_0m + v10 + a_21;
      ws[k][3] = v00 + v11 + a_22;
    }
  }
  __syncthreads();
}

template <int S, int POOL, bool NTK>
__global__ __launch_bounds__(256) void k_stage_fast(
    const float* __restrict__ in_n, const float* __restrict__ in_t, int n1,
    int n2, int ndiag, int ngroups, int nblk, float wv9, float b_var,
    const float* __restrict__ q1, const float* __restrict__ q2,
    float* __restrict__ out_n, float* __restrict__ out_t) {
  constexpr int H2 = S / 2;
  constexpr int TPP = H2 * H2;            // threads per plane
  constexpr int JPW = 256 / TPP;          // pairs per workgroup
  constexpr int LS = S + 2;
  constexpr int PL = LS * LS;
  constexpr int NT = NTK ? 2 : 1;
  constexpr int NB = POOL == 2 ? 6 : NT;  // LDS planes per exchange
  constexpr int SO = S / POOL;
  constexpr long long PER_IN = (long long)S * S * S * S;
  constexpr long long PER_OUT = (long long)SO * SO * SO * SO;
  __shared__ float lds[JPW * NB * PL];
  __shared__ float2 lq[JPW * 2 * S * S];

  const int L = xcd_remap(blockIdx.x, nblk);
  if (L >= ndiag * ngroups) return;       // whole block exits together
  const int g = L / ndiag;
  const int di = L - g * ndiag;
  const int job = threadIdx.x / TPP;
  const int t = threadIdx.x - job * TPP;
  const int W = t / H2;
  const int Wp = t - W * H2;
  const bool self = n2 == 0;
  const long long T = self ? (long long)n1 : (long long)n1 * n2;
  const long long p = (long long)g * JPW + job;
  const bool valid = p < T;
  const long long pc = valid ? p : 0;
  const long long i = self ? pc : pc / n2;
  const long long j = self ? pc : pc - i * n2;

  for (int k = threadIdx.x; k < JPW * NB * PL; k += 256) lds[k] = 0.0f;
  float2* lq1 = lq + job * 2 * S * S;
  float2* lq2 = lq1 + S * S;
  {
    const float* qa = q1 + i * (S * S);
    const float* qb = (self ? q1 : q2) + j * (S * S);
    for (int k = t; k < S * S; k += TPP) {
      lq1[k] = qnorm(valid ? qa[k] : 0.0f);
      lq2[k] = qnorm(valid ? qb[k] : 0.0f);
    }
  }
  __syncthreads();
  float* L0 = lds + job * NB * PL;
  const float* bn = in_n + pc * PER_IN;
  const float* bt = NTK ? in_t + pc * PER_IN : in_n;

  if constexpr (POOL == 2) {
    const int D = di - (H2 - 1);
    const int Hs = D > 0 ? D : 0;
    const int He = D < 0 ? H2 - 1 + D : H2 - 1;
    const int d0 = 2 * D, dp = 2 * D + 1, dm = 2 * D - 1;
    float c0a[NT][4], c0b[NT][4], cp[NT][4], cm[NT][4];
    {
      const int h0 = 2 * Hs;
      const int hh[4] = {h0 - 1, h0, h0, h0 - 1};
      const int hp[4] = {h0 - 1 - d0, h0 - d0, h0 - dp, h0 - 1 - dm};
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) {
        const float* b0 = tt ? bt : bn;
        const float* const base[4] = {b0, b0, b0, b0};
        float ws[4][4];
        wpass<S, 4>(base, hh, hp, L0, W, Wp, valid, ws);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          c0a[tt][e] = ws[0][e];
          c0b[tt][e] = ws[1][e];
          cp[tt][e] = ws[2][e];
          cm[tt][e] = ws[3][e];
        }
      }
    }
    for (int H = Hs; H <= He; ++H) {
      const int Hp = H - D;
      const int hh[6] = {2 * H + 1, 2 * H + 2, 2 * H + 1,
                         2 * H + 2, 2 * H,     2 * H + 1};
      const int hp[6] = {2 * H + 1 - d0, 2 * H + 2 - d0, 2 * H + 1 - dp,
                         2 * H + 2 - dp, 2 * H - dm,     2 * H + 1 - dm};
      float pre[NT][4][4];   // [tensor][P00, P11, P10, P01][element]
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) {
        const float* b0 = tt ? bt : bn;
        const float* const base[6] = {b0, b0, b0, b0, b0, b0};
        float ws[6][4];
        wpass<S, 6>(base, hh, hp, L0, W, Wp, valid, ws);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          pre[tt][0][e] = c0a[tt][e] + c0b[tt][e] + ws[0][e];
          pre[tt][1][e] = c0b[tt][e] + ws[0][e] + ws[1][e];
          pre[tt][2][e] = cp[tt][e] + ws[2][e] + ws[3][e];
          pre[tt][3][e] = cm[tt][e] + ws[4][e] + ws[5][e];
          c0a[tt][e] = ws[0][e];
          c0b[tt][e] = ws[1][e];
          cp[tt][e] = ws[3][e];
          cm[tt][e] = ws[5][e];
        }
      }
      const int ra[4] = {2 * H, 2 * H + 1, 2 * H + 1, 2 * H};
      const int rb[4] = {2 * Hp, 2 * Hp + 1, 2 * Hp, 2 * Hp + 1};
      float acc_n = 0.0f, acc_t = 0.0f;
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float2 qa = lq1[ra[pl] * S + 2 * W + (e >> 1)];
          const float2 qb = lq2[rb[pl] * S + 2 * Wp + (e & 1)];
          const float cv = fmaf(wv9, pre[0][pl][e], b_var);
          float ds;
          acc_n += relu_map_norm(cv, qa.x * qb.x, qa.y * qb.y, &ds);
          if constexpr (NTK) acc_t += fmaf(wv9, pre[NT - 1][pl][e], cv) * ds;
        }
      }
      if (valid) {
        const long long o =
            p * PER_OUT + ((long long)(H * SO + Hp) * SO + W) * SO + Wp;
        out_n[o] = acc_n * 0.0625f;
        if constexpr (NTK) out_t[o] = acc_t * 0.0625f;
      }
    }
  } else {
    const int d = di - (S - 1);
    const int hs = d > 0 ? d : 0;
    const int he = d < 0 ? S - 1 + d : S - 1;
    float cA[NT][4], cB[NT][4];
    {
      int hh[NT], hp[NT];
      const float* base[NT];
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) {
        hh[tt] = hs;
        hp[tt] = hs - d;
        base[tt] = tt ? bt : bn;
      }
      float ws[NT][4];
      wpass<S, NT>(base, hh, hp, L0, W, Wp, valid, ws);
#pragma unroll
      for (int tt = 0; tt < NT; ++tt)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          cA[tt][e] = 0.0f;        // plane (hs-1, hs-1-d) is always outside
          cB[tt][e] = ws[tt][e];
        }
    }
    for (int h = hs; h <= he; ++h) {
      const int hq = h - d;
      int hh[NT], hp[NT];
      const float* base[NT];
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) {
        hh[tt] = h + 1;
        hp[tt] = h + 1 - d;
        base[tt] = tt ? bt : bn;
      }
      float ws[NT][4];
      wpass<S, NT>(base, hh, hp, L0, W, Wp, valid, ws);
      float pre[NT][4];
#pragma unroll
      for (int tt = 0; tt < NT; ++tt)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          pre[tt][e] = cA[tt][e] + cB[tt][e] + ws[tt][e];
          cA[tt][e] = cB[tt][e];
          cB[tt][e] = ws[tt][e];
        }
      float vn[4], vt[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float2 qa = lq1[h * S + 2 * W + (e >> 1)];
        const float2 qb = lq2[hq * S + 2 * Wp + (e & 1)];
        const float cv = fmaf(wv9, pre[0][e], b_var);
        float ds;
        vn[e] = relu_map_norm(cv, qa.x * qb.x, qa.y * qb.y, &ds);
        vt[e] = NTK ? fmaf(wv9, pre[NT - 1][e], cv) * ds : 0.0f;
      }
      if (valid) {
        const long long o = p * PER_OUT + ((long long)(h * S + hq) * S +
                                           2 * W) * S + 2 * Wp;
        *reinterpret_cast<float2*>(out_n + o) = make_float2(vn[0], vn[1]);
        *reinterpret_cast<float2*>(out_n + o + S) = make_float2(vn[2], vn[3]);
        if constexpr (NTK) {
          *reinterpret_cast<float2*>(out_t + o) = make_float2(vt[0], vt[1]);
          *reinterpret_cast<float2*>(out_t + o + S) =
              make_float2(vt[2], vt[3]);
        }
      }
    }
  }
}

// ------------------------------------------------------------ first stage

// Loads image n of x ([N][S][S][C], C <= 4) into a zero-padded,
// channel-planar LDS image xp[4][S+2][S+2] (channels >= C are zero).
template <int S>
__device__ __forceinline__ void load_padded(const float* __restrict__ x,
                                            long long n, int C, bool valid,
                                            float* __restrict__ xp) {
  constexpr int SP = S + 2;
  for (int k = threadIdx.x; k < 4 * SP * SP; k += blockDim.x) {
    const int c = k / (SP * SP);
    const int r = k - c * (SP * SP);
    const int hh = r / SP - 1;
    const int ww = r - (r / SP) * SP - 1;
    float v = 0.0f;
    if (valid && c < C && (unsigned)hh < (unsigned)S &&
        (unsigned)ww < (unsigned)S)
      v = x[((n * S + hh) * S + ww) * C + c];
    xp[k] = v;
  }
}

template <bool NTK>
__global__ __launch_bounds__(256) void k_first32(
    const float* __restrict__ x1, const float* __restrict__ x2, int n1,
    int n2, int C, int nblk, float scale, float b_var,
    const float* __restrict__ q1, const float* __restrict__ q2,
    float* __restrict__ out_n, float* __restrict__ out_t) {
  constexpr int S = 32, SP = S + 2, PLN = SP * SP;
  constexpr long long PER = (long long)S * S * S * S;
  __shared__ float xp1[4 * PLN];
  __shared__ float xp2[4 * PLN];
  __shared__ float2 lq2[S * S];
  __shared__ float2 lq1[S];
  const bool self = n2 == 0;
  const long long T = self ? (long long)n1 : (long long)n1 * n2;
  const int L = xcd_remap(blockIdx.x, nblk);
  if ((long long)L >= T * S) return;
  const long long p = L / S;
  const int h = L - (int)(p * S);
  const long long i = self ? p : p / n2;
  const long long j = self ? p : p - i * n2;
  const float* xb = self ? x1 : x2;
  const float* qbp = self ? q1 : q2;
  load_padded<S>(x1, i, C, true, xp1);
  load_padded<S>(xb, j, C, true, xp2);
  for (int k = threadIdx.x; k < S * S; k += 256)
    lq2[k] = qnorm(qbp[j * (S * S) + k]);
  if (threadIdx.x < S) lq1[threadIdx.x] = qnorm(q1[(i * S + h) * S +
                                                   threadIdx.x]);
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int col = lane & 31;
  const int half = lane >> 5;
  // K = 9 taps x 4 channels; step kk covers tap kk >> 1 and channels
  // 2 (kk & 1) + {0, 1} (lane half).
  float a[18];
  const float* pa = xp1 + half * PLN + h * SP + col;
#pragma unroll
  for (int kk = 0; kk < 18; ++kk) {
    const int tap = kk >> 1;
    a[kk] = pa[2 * (kk & 1) * PLN + (tap / 3) * SP + tap % 3];
  }
  float sq1r[16], rq1r[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int m = 8 * (r >> 2) + 4 * half + (r & 3);
    sq1r[r] = lq1[m].x;
    rq1r[r] = lq1[m].y;
  }
  for (int hq = wave; hq < S; hq += 4) {
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    const float* pb = xp2 + half * PLN + hq * SP + col;
#pragma unroll
    for (int kk = 0; kk < 18; ++kk) {
      const int tap = kk >> 1;
      const float bv = pb[2 * (kk & 1) * PLN + (tap / 3) * SP + tap % 3];
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[kk], bv, acc, 0, 0, 0);
    }
    const float2 qb = lq2[hq * S + col];
    float* on = out_n + p * PER + (long long)(h * S + hq) * (S * S);
    float* ot = NTK ? out_t + p * PER + (long long)(h * S + hq) * (S * S)
                    : nullptr;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int m = 8 * (r >> 2) + 4 * half + (r & 3);
      const float cv = fmaf(scale, acc[r], b_var);
      float ds;
      const float v = relu_map_norm(cv, sq1r[r] * qb.x, rq1r[r] * qb.y, &ds);
      on[m * S + col] = v;
      if constexpr (NTK) ot[m * S + col] = cv * ds;
    }
  }
}

template <bool NTK>
__global__ __launch_bounds__(256) void k_first16(
    const float* __restrict__ x1, const float* __restrict__ x2, int n1,
    int n2, int C, int nblk, float scale, float b_var,
    const float* __restrict__ q1, const float* __restrict__ q2,
    float* __restrict__ out_n, float* __restrict__ out_t) {
  constexpr int S = 16, SP = S + 2, PLN = SP * SP;
  constexpr long long PER = (long long)S * S * S * S;
  __shared__ float xp1[4 * PLN];
  __shared__ float xp2[4 * PLN];
  __shared__ float2 lq2[S * S];
  __shared__ float2 lq1[S];
  const bool self = n2 == 0;
  const long long T = self ? (long long)n1 : (long long)n1 * n2;
  const int L = xcd_remap(blockIdx.x, nblk);
  if ((long long)L >= T * S) return;
  const long long p = L / S;
  const int h = L - (int)(p * S);
  const long long i = self ? p : p / n2;
  const long long j = self ? p : p - i * n2;
  const float* xb = self ? x1 : x2;
  const float* qbp = self ? q1 : q2;
  load_padded<S>(x1, i, C, true, xp1);
  load_padded<S>(xb, j, C, true, xp2);
  for (int k = threadIdx.x; k < S * S; k += 256)
    lq2[k] = qnorm(qbp[j * (S * S) + k]);
  if (threadIdx.x < S) lq1[threadIdx.x] = qnorm(q1[(i * S + h) * S +
                                                   threadIdx.x]);
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int col = lane & 15;
  const int ch = lane >> 4;           // channel of this lane's K slice
  // K = 9 taps x 4 channels; step kk = tap, the 4 channels across lanes.
  float a[9];
  const float* pa = xp1 + ch * PLN + h * SP + col;
#pragma unroll
  for (int kk = 0; kk < 9; ++kk) a[kk] = pa[(kk / 3) * SP + kk % 3];
  float sq1r[4], rq1r[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = 4 * ch + r;
    sq1r[r] = lq1[m].x;
    rq1r[r] = lq1[m].y;
  }
  for (int hq = wave; hq < S; hq += 4) {
    floatx4 acc;
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = 0.0f;
    const float* pb = xp2 + ch * PLN + hq * SP + col;
#pragma unroll
    for (int kk = 0; kk < 9; ++kk) {
      const float bv = pb[(kk / 3) * SP + kk % 3];
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[kk], bv, acc, 0, 0, 0);
    }
    const float2 qb = lq2[hq * S + col];
    float* on = out_n + p * PER + (long long)(h * S + hq) * (S * S);
    float* ot = NTK ? out_t + p * PER + (long long)(h * S + hq) * (S * S)
                    : nullptr;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = 4 * ch + r;
      const float cv = fmaf(scale, acc[r], b_var);
      float ds;
      const float v = relu_map_norm(cv, sq1r[r] * qb.x, rq1r[r] * qb.y, &ds);
      on[m * S + col] = v;
      if constexpr (NTK) ot[m * S + col] = cv * ds;
    }
  }
}

// ---------------------------------------------------------------- launchers

static inline int pad8(long long n) { return (int)((n + 7) / 8 * 8); }

template <int S, int POOL>
static int launch_stage_s(const float* in_n, const float* in_t, int n1,
                          int n2, bool ntk, float wv9, float b_var,
                          const float* q1, const float* q2, float* out_n,
                          float* out_t, hipStream_t st) {
  constexpr int JPW = 256 / ((S / 2) * (S / 2));
  const long long T = n2 == 0 ? (long long)n1 : (long long)n1 * n2;
  const int ndiag = POOL == 2 ? S - 1 : 2 * S - 1;
  const long long ngroups = (T + JPW - 1) / JPW;
  if (ngroups * ndiag > 0x7ffffff0LL)
    return set_error(NTA_EINVAL, "nta_stage: tile too large (%lld pairs)", T);
  const int nblk = pad8(ngroups * ndiag);
  if (ntk)
    hipLaunchKernelGGL((k_stage_fast<S, POOL, true>), dim3(nblk), dim3(256),
                       0, st, in_n, in_t, n1, n2, ndiag, (int)ngroups, nblk,
                       wv9, b_var, q1, q2, out_n, out_t);
  else
    hipLaunchKernelGGL((k_stage_fast<S, POOL, false>), dim3(nblk), dim3(256),
                       0, st, in_n, in_t, n1, n2, ndiag, (int)ngroups, nblk,
                       wv9, b_var, q1, q2, out_n, out_t);
  return check_launch("nta_stage");
}

int launch_stage_fast(const float* in_nngp, const float* in_ntk, int n1,
                      int n2, int S, int pool, int flags, float w_var,
                      float b_var, const float* q1, const float* q2,
                      float* out_nngp, float* out_ntk, hipStream_t stream) {
  const bool ntk = flags & NTA_NTK;
  const float wv9 = w_var / 9.0f;
  if (pool == 2) {
    switch (S) {
      case 32: return launch_stage_s<32, 2>(in_nngp, in_ntk, n1, n2, ntk, wv9,
                                            b_var, q1, q2, out_nngp, out_ntk,
                                            stream);
      case 16: return launch_stage_s<16, 2>(in_nngp, in_ntk, n1, n2, ntk, wv9,
                                            b_var, q1, q2, out_nngp, out_ntk,
                                            stream);
      case 8: return launch_stage_s<8, 2>(in_nngp, in_ntk, n1, n2, ntk, wv9,
                                          b_var, q1, q2, out_nngp, out_ntk,
                                          stream);
    }
  } else if (pool == 1) {
    switch (S) {
      case 32: return launch_stage_s<32, 1>(in_nngp, in_ntk, n1, n2, ntk, wv9,
                                            b_var, q1, q2, out_nngp, out_ntk,
                                            stream);
      case 16: return launch_stage_s<16, 1>(in_nngp, in_ntk, n1, n2, ntk, wv9,
                                            b_var, q1, q2, out_nngp, out_ntk,
                                            stream);
      case 8: return launch_stage_s<8, 1>(in_nngp, in_ntk, n1, n2, ntk, wv9,
                                          b_var, q1, q2, out_nngp, out_ntk,
                                          stream);
    }
  }
  return kNotHandled;
}

int launch_first_fast(const float* x1, const float* x2, int n1, int n2, int S,
                      int C, int flags, float w_var, float b_var,
                      const float* q1, const float* q2, float* nngp,
                      float* ntk, hipStream_t stream) {
  if (C > 4 || (S != 32 && S != 16)) return kNotHandled;
  const long long T = n2 == 0 ? (long long)n1 : (long long)n1 * n2;
  if (T * S > 0x7ffffff0LL)
    return set_error(NTA_EINVAL, "nta_first_stage: tile too large (%lld "
                     "pairs)", T);
  const int nblk = pad8(T * S);
  const float scale = w_var / (9.0f * (float)C);
  const bool want_ntk = flags & NTA_NTK;
  if (S == 32) {
    if (want_ntk)
      hipLaunchKernelGGL(k_first32<true>, dim3(nblk), dim3(256), 0, stream,
                         x1, x2, n1, n2, C, nblk, scale, b_var, q1, q2, nngp,
                         ntk);
    else
      hipLaunchKernelGGL(k_first32<false>, dim3(nblk), dim3(256), 0, stream,
                         x1, x2, n1, n2, C, nblk, scale, b_var, q1, q2, nngp,
                         ntk);
  } else {
    if (want_ntk)
      hipLaunchKernelGGL(k_first16<true>, dim3(nblk), dim3(256), 0, stream,
                         x1, x2, n1, n2, C, nblk, scale, b_var, q1, q2, nngp,
                         ntk);
    else
      hipLaunchKernelGGL(k_first16<false>, dim3(nblk), dim3(256), 0, stream,
                         x1, x2, n1, n2, C, nblk, scale, b_var, q1, q2, nngp,
                         ntk);
  }
  return check_launch("nta_first_stage");
}

}  // namespace nta
